// Wide forward kernels (one color of every instance per launch, optional tape) and the
// hand-written reverse-mode kernels of the XPBD substep (gfx950).
//
// Reverse formulas (inputs read from the tape, adjoints updated in place; within one
// color no two elements share a vertex, so the read-modify-write of gx is race free):
//
// stretch  d = a - b, l = |d|, D = w1 + w2 + at, at = c / dt^2,
//          dl = (-(l - r) - at lam) / D, a' = a + w1 dl d / (l + eps), b' = b - w2 ...,
//          lam' = lam + dl.  With u = w1 ga' - w2 gb', s = dl / (l + eps):
//            gs = u.d, gd = s u, gdl = glam' + gs / (l + eps), ginv = gs dl,
//            glam = glam' - gdl at / D, gl = -gdl / D - ginv / (l + eps)^2,
//            gat = -gdl (lam + dl) / D, gd += gl d / l, ga = ga' + gd, gb = gb' - gd
// bend     lam = sgn s (ang - th) / denom with d = clamp(n1.n2), s = sqrt(1 - d^2),
//          ang = acos(d); the q_k are differentiated through the cross products
//          (c = a x b: ga += b x gc, gb += gc x a) and the normalisations; dP/dd for
//          P = s (ang - th) is -d (ang - th) / s - 1, taken as 0 at a fold (s = 0) where
//          dd/dx itself vanishes
// ground   free vertex below the plane: y' = h; kinetic friction x'_t = x_t - mu cl t^
//          (t = x_t - x0_t, cl = h - y), static friction x'_t = x0_t
// predict  x = x0 + v' dt            finalize  v = (x - x0) f, x0 = x
#include "adjoint.h"
#include "elements.h"

namespace xrt {

namespace {
constexpr int kBs = 256;
inline int Blocks(int n) { return (n + kBs - 1) / kBs; }

__device__ __forceinline__ float3 F3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 Neg(float3 a) { return make_float3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ void Acc(float3& a, float3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
}  // namespace

// ------------------------------------------------------------------ forward
// (instance, block within its color range) of this block: from the color's block map
// (walked in XCD-contiguous order: the hardware deals workgroups to the 8 XCDs in turn),
// or grid.y / grid.x.  false: idle block.
__device__ __forceinline__ bool WideBlock(const int32_t* __restrict__ map, int n_inst, int& inst,
                                          int& blk) {
  if (!map) {
    inst = blockIdx.y;
    blk = blockIdx.x;
    return true;
  }
  const int total = map[n_inst];
  const int l = static_cast<int>(blockIdx.x & 7u) * ((total + 7) >> 3) + static_cast<int>(blockIdx.x >> 3);
  if (l >= total) return false;
  inst = map[n_inst + 1 + l];
  blk = l - map[inst];
  return true;
}

__global__ __launch_bounds__(kBs) void WideStretchKernel(WideArgs a, int color, float inv_dt2,
                                                         int iter, float4* __restrict__ tape,
                                                         const int32_t* __restrict__ map,
                                                         int n_inst) {
  int inst, blk;
  if (!WideBlock(map, n_inst, inst, blk)) return;
  const InstDesc d = a.inst[inst];
  if (color >= d.s_ncol) return;
  const int beg = a.s_coloff[d.s_cbeg + color], end = a.s_coloff[d.s_cbeg + color + 1];
  const int k = beg + blk * kBs + threadIdx.x;
  if (k >= end) return;
  float4* X = a.x + d.vbeg;
  const StretchC c = a.sc[k];
  if (tape) {
    const float4 xa = X[c.i0];
    tape[2 * k] = make_float4(xa.x, xa.y, xa.z, iter == 0 ? 0.0f : a.slam[k]);
    tape[2 * k + 1] = X[c.i1];
  }
  StretchProject(X, c, a.slam + k, inv_dt2, iter);
}

__global__ __launch_bounds__(kBs) void WideBendKernel(WideArgs a, int color, float inv_dt2,
                                                      BendMode m, float4* __restrict__ tape,
                                                      const int32_t* __restrict__ map,
                                                      int n_inst) {
  int inst, blk;
  if (!WideBlock(map, n_inst, inst, blk)) return;
  const InstDesc d = a.inst[inst];
  if (color >= d.b_ncol) return;
  const int beg = a.b_coloff[d.b_cbeg + color], end = a.b_coloff[d.b_cbeg + color + 1];
  const int k = beg + blk * kBs + threadIdx.x;
  if (k >= end) return;
  float4* X = a.x + d.vbeg;
  const int4 id = a.bidx[k];
  if (tape) {
    const float4 x1 = X[id.x], x2 = X[id.y], x3 = X[id.z], x4 = X[id.w];
    tape[3 * k] = make_float4(x1.x, x1.y, x1.z, x2.x);
    tape[3 * k + 1] = make_float4(x2.y, x2.z, x3.x, x3.y);
    tape[3 * k + 2] = make_float4(x3.z, x4.x, x4.y, x4.z);
  }
  BendProject(X, id, a.bpar[k], inv_dt2, m);
}

// The same launches with two elements per thread (kBs / 2 threads per block, the block
// maps unchanged): both elements' records and vertex gathers are issued before either is
// solved.  Elements of one color share no vertex, so the order cannot change a bit.
__global__ __launch_bounds__(kBs / 2) void WideStretch2Kernel(WideArgs a, int color,
                                                              float inv_dt2, int iter,
                                                              const int32_t* __restrict__ map,
                                                              int n_inst) {
  int inst, blk;
  if (!WideBlock(map, n_inst, inst, blk)) return;
  const InstDesc d = a.inst[inst];
  if (color >= d.s_ncol) return;
  const int beg = a.s_coloff[d.s_cbeg + color] + blk * kBs;
  const int end = min(a.s_coloff[d.s_cbeg + color + 1], beg + kBs);
  StretchColor<2>(a.x + d.vbeg, a.sc, a.slam, beg, end, inv_dt2, iter, static_cast<int>(threadIdx.x),
                  kBs / 2);
}

__global__ __launch_bounds__(kBs / 2) void WideBend2Kernel(WideArgs a, int color, float inv_dt2,
                                                           BendMode m,
                                                           const int32_t* __restrict__ map,
                                                           int n_inst) {
  int inst, blk;
  if (!WideBlock(map, n_inst, inst, blk)) return;
  const InstDesc d = a.inst[inst];
  if (color >= d.b_ncol) return;
  const int beg = a.b_coloff[d.b_cbeg + color] + blk * kBs;
  const int end = min(a.b_coloff[d.b_cbeg + color + 1], beg + kBs);
  BendColor<2>(a.x + d.vbeg, a.bidx, a.bpar, beg, end, inv_dt2, m, static_cast<int>(threadIdx.x),
               kBs / 2);
}

__global__ __launch_bounds__(kBs) void Copy4Kernel(float4* __restrict__ dst,
                                                   const float4* __restrict__ src, int n) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ __launch_bounds__(kBs) void Zero4Kernel(float4* __restrict__ p, int n) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__global__ __launch_bounds__(kBs) void Zero1Kernel(float* __restrict__ p, int n) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// ------------------------------------------------------------------ reverse
__global__ __launch_bounds__(kBs) void FinalizeAdjKernel(float4* __restrict__ gx0,
                                                         const float4* __restrict__ gv,
                                                         float4* __restrict__ gx, int n, float f) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  const float4 a = gx0[i], b = gv[i];
  gx[i] = make_float4(a.x + b.x * f, a.y + b.y * f, a.z + b.z * f, 0.0f);
  gx0[i] = make_float4(-b.x * f, -b.y * f, -b.z * f, 0.0f);
}

__global__ __launch_bounds__(kBs) void PredictAdjKernel(float4* __restrict__ gx0,
                                                        float4* __restrict__ gv,
                                                        const float4* __restrict__ gx, int n,
                                                        float dt) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  const float4 g = gx[i], a = gx0[i];
  gx0[i] = make_float4(a.x + g.x, a.y + g.y, a.z + g.z, 0.0f);
  gv[i] = make_float4(g.x * dt, g.y * dt, g.z * dt, 0.0f);
}

__global__ __launch_bounds__(kBs) void GroundAdjKernel(const float4* __restrict__ xpre,
                                                       const float4* __restrict__ x0s,
                                                       float4* __restrict__ gx,
                                                       float4* __restrict__ gx0, int n,
                                                       float ground_y, float mu) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  const float4 x = xpre[i];
  if (x.w <= kEps || x.y >= ground_y) return;  // the plane did not act: identity
  const float4 g = gx[i];
  const float cl = ground_y - x.y;
  float4 out = make_float4(g.x, 0.0f, g.z, 0.0f);  // y' = ground_y
  if (mu > 0.0f && cl > 0.0f) {
    const float4 p0 = x0s[i];
    const float tx = x.x - p0.x, tz = x.z - p0.z;
    const float tl = sqrtf(tx * tx + tz * tz);
    const float mx = cl * mu;
    const float r = mx / tl;
    float4 a0 = gx0[i];
    if (tl > 0.0f && r < 1.0f) {  // kinetic: x'_t = x_t - mx t/|t|
      const float hx = tx / tl, hz = tz / tl;
      const float gt = hx * g.x + hz * g.z;
      const float bx = -r * (g.x - hx * gt), bz = -r * (g.z - hz * gt);  // adjoint of t
      out = make_float4(g.x + bx, mu * gt, g.z + bz, 0.0f);
      a0.x -= bx;
      a0.z -= bz;
    } else {  // static: x'_t = x0_t
      out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      a0.x += g.x;
      a0.z += g.z;
    }
    gx0[i] = a0;
  }
  gx[i] = out;
}

__global__ __launch_bounds__(kBs) void WideStretchAdjKernel(WideArgs a, int color, float inv_dt2,
                                                            int iter,
                                                            const float4* __restrict__ tape,
                                                            float4* __restrict__ gx,
                                                            float* __restrict__ glam,
                                                            float* __restrict__ gcomp) {
  const InstDesc di = a.inst[blockIdx.y];
  if (color >= di.s_ncol) return;
  const int beg = a.s_coloff[di.s_cbeg + color], end = a.s_coloff[di.s_cbeg + color + 1];
  const int k = beg + blockIdx.x * kBs + threadIdx.x;
  if (k >= end) return;
  const StretchC c = a.sc[k];
  const int i0 = di.vbeg + c.i0, i1 = di.vbeg + c.i1;
  const float4 ta = tape[2 * k], tb = tape[2 * k + 1];
  const float lam = ta.w;
  const float w1 = a.xw[i0].w, w2 = a.xw[i1].w;
  const float dx = ta.x - tb.x, dy = ta.y - tb.y, dz = ta.z - tb.z;
  const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
  float gl = glam[k];
  if (dist != c.rest && w1 + w2 > 0.0f) {
    const float4 ga = gx[i0], gb = gx[i1];
    const float inv = 1.0f / (dist + kEps);
    const float at = c.compliance * inv_dt2;
    const float D = w1 + w2 + at;
    const float dl = (-(dist - c.rest) - at * lam) / D;
    const float s = dl * inv;
    const float ux = w1 * ga.x - w2 * gb.x, uy = w1 * ga.y - w2 * gb.y, uz = w1 * ga.z - w2 * gb.z;
    const float gs = ux * dx + uy * dy + uz * dz;
    const float gdl = gl + gs * inv;
    const float ginv = gs * dl;
    const float gdist = -gdl / D - ginv * inv * inv;
    gcomp[k] += -gdl * (lam + dl) / D * inv_dt2;
    gl = gl - gdl * at / D;
    const float gd = dist > 0.0f ? gdist / dist : 0.0f;
    const float ex = s * ux + gd * dx, ey = s * uy + gd * dy, ez = s * uz + gd * dz;
    gx[i0] = make_float4(ga.x + ex, ga.y + ey, ga.z + ez, 0.0f);
    gx[i1] = make_float4(gb.x - ex, gb.y - ey, gb.z - ez, 0.0f);
  }
  glam[k] = iter == 0 ? 0.0f : gl;
}

__global__ __launch_bounds__(kBs) void WideBendAdjKernel(WideArgs a, int color, float inv_dt2,
                                                         BendMode m,
                                                         const float4* __restrict__ tape,
                                                         float4* __restrict__ gx,
                                                         float* __restrict__ gbcomp) {
  const InstDesc di = a.inst[blockIdx.y];
  if (color >= di.b_ncol) return;
  const int beg = a.b_coloff[di.b_cbeg + color], end = a.b_coloff[di.b_cbeg + color + 1];
  const int k = beg + blockIdx.x * kBs + threadIdx.x;
  if (k >= end) return;
  const int4 lid = a.bidx[k];
  const int id1 = di.vbeg + lid.x, id2 = di.vbeg + lid.y, id3 = di.vbeg + lid.z, id4 = di.vbeg + lid.w;
  const float4 t0 = tape[3 * k], t1 = tape[3 * k + 1], t2 = tape[3 * k + 2];
  const float3 x1 = make_float3(t0.x, t0.y, t0.z), x2 = make_float3(t0.w, t1.x, t1.y);
  const float3 x3 = make_float3(t1.z, t1.w, t2.x), x4 = make_float3(t2.y, t2.z, t2.w);
  const float w1 = a.xw[id1].w, w2 = a.xw[id2].w, w3 = a.xw[id3].w, w4 = a.xw[id4].w;
  const float2 par = a.bpar[k];

  // forward recomputation (BendProject)
  const float3 p2 = Sub(x2, x1), p3 = Sub(x3, x1), p4 = Sub(x4, x1);
  const float3 c23 = Cross(p2, p3), c24 = Cross(p2, p4);
  const float l23r = sqrtf(Dot(c23, c23)), l24r = sqrtf(Dot(c24, c24));
  const float3 n1 = Scale(c23, 1.0f / l23r), n2 = Scale(c24, 1.0f / l24r);
  const float draw = Dot(n1, n2);
  if (draw != draw) return;
  const float d = fminf(fmaxf(draw, m.dlo), 1.0f);
  const float ang = acosf(d);
  if (ang < kEps) return;
  const float i23 = 1.0f / (l23r + kEps), i24 = 1.0f / (l24r + kEps);
  const float3 A3 = Add(Cross(p2, n2), Scale(Cross(n1, p2), d));
  const float3 A4 = Add(Cross(p2, n1), Scale(Cross(n2, p2), d));
  const float3 A2a = Add(Cross(p3, n2), Scale(Cross(n1, p3), d));
  const float3 A2b = Add(Cross(p4, n1), Scale(Cross(n2, p4), d));
  const float3 q3 = Scale(A3, i23), q4 = Scale(A4, i24);
  const float3 q2 = Sub(Scale(A2a, -i23), Scale(A2b, i24));
  const float3 q1 = Scale(Add(Add(q2, q3), q4), -1.0f);
  const float denom = par.y * inv_dt2 +
                      (w1 * Dot(q1, q1) + w2 * Dot(q2, q2) + w3 * Dot(q3, q3) + w4 * Dot(q4, q4));
  if (denom < kEps) return;
  const float th = m.th0 + m.th1 * par.x;
  const float s = sqrtf(1.0f - d * d);
  const float lam = m.sgn * s * (ang - th) / denom;

  // reverse
  const float3 g1 = F3(gx[id1]), g2 = F3(gx[id2]), g3 = F3(gx[id3]), g4 = F3(gx[id4]);
  const float glam = w1 * Dot(q1, g1) + w2 * Dot(q2, g2) + w3 * Dot(q3, g3) + w4 * Dot(q4, g4);
  const float gden = -glam * lam / denom;
  gbcomp[k] += gden * inv_dt2;
  float3 gq1 = Add(Scale(g1, w1 * lam), Scale(q1, 2.0f * gden * w1));
  float3 gq2 = Add(Scale(g2, w2 * lam), Scale(q2, 2.0f * gden * w2));
  float3 gq3 = Add(Scale(g3, w3 * lam), Scale(q3, 2.0f * gden * w3));
  float3 gq4 = Add(Scale(g4, w4 * lam), Scale(q4, 2.0f * gden * w4));
  const float gP = glam * m.sgn / denom;
  float gd = s > 0.0f ? gP * (-d * (ang - th) / s - 1.0f) : 0.0f;
  // q1 = -(q2 + q3 + q4)
  Acc(gq2, Neg(gq1));
  Acc(gq3, Neg(gq1));
  Acc(gq4, Neg(gq1));
  // q3 = A3 i23, q4 = A4 i24, q2 = -A2a i23 - A2b i24
  const float3 gA3 = Scale(gq3, i23), gA4 = Scale(gq4, i24);
  const float3 gA2a = Scale(gq2, -i23), gA2b = Scale(gq2, -i24);
  const float gi23 = Dot(gq3, A3) - Dot(gq2, A2a);
  const float gi24 = Dot(gq4, A4) - Dot(gq2, A2b);
  float3 gp2 = make_float3(0.0f, 0.0f, 0.0f), gp3 = gp2, gp4 = gp2, gn1 = gp2, gn2 = gp2;
  // A3 = p2 x n2 + d (n1 x p2)
  Acc(gp2, Cross(n2, gA3)); Acc(gn2, Cross(gA3, p2));
  gd += Dot(gA3, Cross(n1, p2));
  { const float3 t = Scale(gA3, d); Acc(gn1, Cross(p2, t)); Acc(gp2, Cross(t, n1)); }
  // A4 = p2 x n1 + d (n2 x p2)
  Acc(gp2, Cross(n1, gA4)); Acc(gn1, Cross(gA4, p2));
  gd += Dot(gA4, Cross(n2, p2));
  { const float3 t = Scale(gA4, d); Acc(gn2, Cross(p2, t)); Acc(gp2, Cross(t, n2)); }
  // A2a = p3 x n2 + d (n1 x p3)
  Acc(gp3, Cross(n2, gA2a)); Acc(gn2, Cross(gA2a, p3));
  gd += Dot(gA2a, Cross(n1, p3));
  { const float3 t = Scale(gA2a, d); Acc(gn1, Cross(p3, t)); Acc(gp3, Cross(t, n1)); }
  // A2b = p4 x n1 + d (n2 x p4)
  Acc(gp4, Cross(n1, gA2b)); Acc(gn1, Cross(gA2b, p4));
  gd += Dot(gA2b, Cross(n2, p4));
  { const float3 t = Scale(gA2b, d); Acc(gn2, Cross(p4, t)); Acc(gp4, Cross(t, n2)); }
  // i23 = 1 / (|c23| + eps): d|c23|/dc23 = n1
  const float gl23 = -gi23 * i23 * i23, gl24 = -gi24 * i24 * i24;
  // clamp, then d = n1 . n2
  const float gdr = (draw < m.dlo || draw > 1.0f) ? 0.0f : gd;
  Acc(gn1, Scale(n2, gdr));
  Acc(gn2, Scale(n1, gdr));
  // n = c / |c|: gc = (gn - n (n . gn)) / |c| + g|c| n
  const float3 gc23 = Add(Scale(Sub(gn1, Scale(n1, Dot(n1, gn1))), 1.0f / l23r), Scale(n1, gl23));
  const float3 gc24 = Add(Scale(Sub(gn2, Scale(n2, Dot(n2, gn2))), 1.0f / l24r), Scale(n2, gl24));
  // c23 = p2 x p3, c24 = p2 x p4
  Acc(gp2, Cross(p3, gc23)); Acc(gp3, Cross(gc23, p2));
  Acc(gp2, Cross(p4, gc24)); Acc(gp4, Cross(gc24, p2));
  // p_k = x_k - x1, plus the identity part of x' = x + w lam q
  const float3 o1 = Sub(g1, Add(Add(gp2, gp3), gp4));
  const float3 o2 = Add(g2, gp2), o3 = Add(g3, gp3), o4 = Add(g4, gp4);
  gx[id1] = make_float4(o1.x, o1.y, o1.z, 0.0f);
  gx[id2] = make_float4(o2.x, o2.y, o2.z, 0.0f);
  gx[id3] = make_float4(o3.x, o3.y, o3.z, 0.0f);
  gx[id4] = make_float4(o4.x, o4.y, o4.z, 0.0f);
}

// ------------------------------------------------------------------ launchers
hipError_t LaunchWideStretch(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                             float inv_dt2, int iter, float4* tape, const int32_t* map,
                             int blocks, int per_thread) {
  if (max_cnt <= 0 || n_inst <= 0) return hipSuccess;
  if (map && !tape && per_thread == 2) {
    if (blocks > 0)
      WideStretch2Kernel<<<(blocks + 7) / 8 * 8, kBs / 2, 0, s>>>(a, color, inv_dt2, iter, map,
                                                                   n_inst);
  } else if (map) {
    if (blocks > 0)
      WideStretchKernel<<<(blocks + 7) / 8 * 8, kBs, 0, s>>>(a, color, inv_dt2, iter, tape, map, n_inst);
  } else {
    WideStretchKernel<<<dim3(Blocks(max_cnt), n_inst), kBs, 0, s>>>(a, color, inv_dt2, iter, tape,
                                                                    nullptr, n_inst);
  }
  return hipGetLastError();
}
hipError_t LaunchWideBend(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                          float inv_dt2, const BendMode& m, float4* tape, const int32_t* map,
                          int blocks, int per_thread) {
  if (max_cnt <= 0 || n_inst <= 0) return hipSuccess;
  if (map && !tape && per_thread == 2) {
    if (blocks > 0)
      WideBend2Kernel<<<(blocks + 7) / 8 * 8, kBs / 2, 0, s>>>(a, color, inv_dt2, m, map, n_inst);
  } else if (map) {
    if (blocks > 0)
      WideBendKernel<<<(blocks + 7) / 8 * 8, kBs, 0, s>>>(a, color, inv_dt2, m, tape, map, n_inst);
  } else {
    WideBendKernel<<<dim3(Blocks(max_cnt), n_inst), kBs, 0, s>>>(a, color, inv_dt2, m, tape,
                                                                 nullptr, n_inst);
  }
  return hipGetLastError();
}
hipError_t LaunchCopy4(hipStream_t s, float4* dst, const float4* src, int n) {
  if (n <= 0) return hipSuccess;
  Copy4Kernel<<<Blocks(n), kBs, 0, s>>>(dst, src, n);
  return hipGetLastError();
}
hipError_t LaunchZero4(hipStream_t s, float4* p, int n) {
  if (n <= 0) return hipSuccess;
  Zero4Kernel<<<Blocks(n), kBs, 0, s>>>(p, n);
  return hipGetLastError();
}
hipError_t LaunchZero1(hipStream_t s, float* p, int n) {
  if (n <= 0) return hipSuccess;
  Zero1Kernel<<<Blocks(n), kBs, 0, s>>>(p, n);
  return hipGetLastError();
}
hipError_t LaunchFinalizeAdj(hipStream_t s, float4* gx0, const float4* gv, float4* gx, int n,
                             float f) {
  if (n <= 0) return hipSuccess;
  FinalizeAdjKernel<<<Blocks(n), kBs, 0, s>>>(gx0, gv, gx, n, f);
  return hipGetLastError();
}
hipError_t LaunchPredictAdj(hipStream_t s, float4* gx0, float4* gv, const float4* gx, int n,
                            float dt) {
  if (n <= 0) return hipSuccess;
  PredictAdjKernel<<<Blocks(n), kBs, 0, s>>>(gx0, gv, gx, n, dt);
  return hipGetLastError();
}
hipError_t LaunchGroundAdj(hipStream_t s, const float4* xpre, const float4* x0s, float4* gx,
                           float4* gx0, int n, float ground_y, float mu) {
  if (n <= 0) return hipSuccess;
  GroundAdjKernel<<<Blocks(n), kBs, 0, s>>>(xpre, x0s, gx, gx0, n, ground_y, mu);
  return hipGetLastError();
}
hipError_t LaunchWideStretchAdj(hipStream_t s, const WideArgs& a, int n_inst, int color,
                                int max_cnt, float inv_dt2, int iter, const float4* tape,
                                float4* gx, float* glam, float* gcomp) {
  if (max_cnt <= 0 || n_inst <= 0) return hipSuccess;
  WideStretchAdjKernel<<<dim3(Blocks(max_cnt), n_inst), kBs, 0, s>>>(a, color, inv_dt2, iter, tape,
                                                                      gx, glam, gcomp);
  return hipGetLastError();
}
hipError_t LaunchWideBendAdj(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                             float inv_dt2, const BendMode& m, const float4* tape, float4* gx,
                             float* gbcomp) {
  if (max_cnt <= 0 || n_inst <= 0) return hipSuccess;
  WideBendAdjKernel<<<dim3(Blocks(max_cnt), n_inst), kBs, 0, s>>>(a, color, inv_dt2, m, tape, gx,
                                                                   gbcomp);
  return hipGetLastError();
}

}  // namespace xrt
