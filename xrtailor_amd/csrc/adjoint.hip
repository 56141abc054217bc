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
// The reverse runs in FP64: the adjoints (gx, gx0, gv as double4, the multiplier and
// compliance adjoints as double) and every derivative; the forward quantities come from the
// FP32 tape.  Each kernel first repeats the forward's FP32 activity tests on the taped
// input (same expressions, no FMA contraction), so the active set is the forward's, then
// differentiates in double.  With stiff compliances the per-iteration terms of the
// compliance and multiplier adjoints nearly cancel; FP32 there loses percents.
namespace {
struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(float3 a) { return D3{a.x, a.y, a.z}; }
__device__ __forceinline__ D3 d3(float4 a) { return D3{a.x, a.y, a.z}; }
__device__ __forceinline__ D3 d3(double4 a) { return D3{a.x, a.y, a.z}; }
__device__ __forceinline__ double4 d4(D3 a) { return make_double4(a.x, a.y, a.z, 0.0); }
__device__ __forceinline__ D3 Cr(D3 a, D3 b) {
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double Dt(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 Sc(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 Ad(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 Sb(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ void Ac(D3& a, D3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
}  // namespace

__global__ __launch_bounds__(kBs) void FinalizeAdjKernel(double4* __restrict__ gx0,
                                                         const double4* __restrict__ gv,
                                                         double4* __restrict__ gx, int n, double f) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  const double4 a = gx0[i], b = gv[i];
  gx[i] = make_double4(a.x + b.x * f, a.y + b.y * f, a.z + b.z * f, 0.0);
  gx0[i] = make_double4(-b.x * f, -b.y * f, -b.z * f, 0.0);
}

__global__ __launch_bounds__(kBs) void PredictAdjKernel(double4* __restrict__ gx0,
                                                        double4* __restrict__ gv,
                                                        const double4* __restrict__ gx, int n,
                                                        double dt) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  const double4 g = gx[i], a = gx0[i];
  gx0[i] = make_double4(a.x + g.x, a.y + g.y, a.z + g.z, 0.0);
  gv[i] = make_double4(g.x * dt, g.y * dt, g.z * dt, 0.0);
}

__global__ __launch_bounds__(kBs) void GroundAdjKernel(const float4* __restrict__ xpre,
                                                       const float4* __restrict__ x0s,
                                                       double4* __restrict__ gx,
                                                       double4* __restrict__ gx0, int n,
                                                       float ground_y, float mu) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i >= n) return;
  const float4 x = xpre[i];
  if (x.w <= kEps || x.y >= ground_y) return;  // the plane did not act: identity
  const double4 g = gx[i];
  double4 out = make_double4(g.x, 0.0, g.z, 0.0);  // y' = ground_y
  // the forward's friction branch (GroundVertex, FP32)
  const float cl = ground_y - x.y;
  if (mu > 0.0f && cl > 0.0f) {
    const float4 p0 = x0s[i];
    const float tx = x.x - p0.x, tz = x.z - p0.z;
    const float tl = sqrtf(tx * tx + tz * tz);
    const float mx = cl * mu;
    double4 a0 = gx0[i];
    if (tl > 0.0f && mx / tl < 1.0f) {  // kinetic: x'_t = x_t - mx t/|t|
      const double dtx = static_cast<double>(x.x) - p0.x, dtz = static_cast<double>(x.z) - p0.z;
      const double dtl = sqrt(dtx * dtx + dtz * dtz);
      const double r = (static_cast<double>(ground_y) - x.y) * mu / dtl;
      const double hx = dtx / dtl, hz = dtz / dtl;
      const double gt = hx * g.x + hz * g.z;
      const double bx = -r * (g.x - hx * gt), bz = -r * (g.z - hz * gt);  // adjoint of t
      out = make_double4(g.x + bx, static_cast<double>(mu) * gt, g.z + bz, 0.0);
      a0.x -= bx;
      a0.z -= bz;
    } else {  // static: x'_t = x0_t
      out = make_double4(0.0, 0.0, 0.0, 0.0);
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
                                                            double4* __restrict__ gx,
                                                            double* __restrict__ glam,
                                                            double* __restrict__ gcomp) {
#pragma clang fp contract(off)
  const InstDesc di = a.inst[blockIdx.y];
  if (color >= di.s_ncol) return;
  const int beg = a.s_coloff[di.s_cbeg + color], end = a.s_coloff[di.s_cbeg + color + 1];
  const int k = beg + blockIdx.x * kBs + threadIdx.x;
  if (k >= end) return;
  const StretchC c = a.sc[k];
  const int i0 = di.vbeg + c.i0, i1 = di.vbeg + c.i1;
  const float4 ta = tape[2 * k], tb = tape[2 * k + 1];
  const float w1 = a.xw[i0].w, w2 = a.xw[i1].w;
  double gl = glam[k];
  {  // the forward's test (StretchSolve, FP32)
    const float fx = ta.x - tb.x, fy = ta.y - tb.y, fz = ta.z - tb.z;
    const float fdist = sqrtf(fx * fx + fy * fy + fz * fz);
    if (!(fdist != c.rest && w1 + w2 > 0.0f)) {
      glam[k] = iter == 0 ? 0.0 : gl;
      return;
    }
  }
  const double lam = ta.w;
  const double dx = static_cast<double>(ta.x) - tb.x, dy = static_cast<double>(ta.y) - tb.y,
               dz = static_cast<double>(ta.z) - tb.z;
  const double dist = sqrt(dx * dx + dy * dy + dz * dz);
  const double4 ga = gx[i0], gb = gx[i1];
  const double inv = 1.0 / (dist + static_cast<double>(kEps));
  const double at = static_cast<double>(c.compliance) * inv_dt2;
  const double D = static_cast<double>(w1) + w2 + at;
  const double dl = (-(dist - c.rest) - at * lam) / D;
  const double s = dl * inv;
  const double ux = w1 * ga.x - w2 * gb.x, uy = w1 * ga.y - w2 * gb.y, uz = w1 * ga.z - w2 * gb.z;
  const double gs = ux * dx + uy * dy + uz * dz;
  const double gdl = gl + gs * inv;
  const double ginv = gs * dl;
  const double gdist = -gdl / D - ginv * inv * inv;
  gcomp[k] += -gdl * (lam + dl) / D * inv_dt2;
  gl = gl - gdl * at / D;
  const double gd = dist > 0.0 ? gdist / dist : 0.0;
  const double ex = s * ux + gd * dx, ey = s * uy + gd * dy, ez = s * uz + gd * dz;
  gx[i0] = make_double4(ga.x + ex, ga.y + ey, ga.z + ez, 0.0);
  gx[i1] = make_double4(gb.x - ex, gb.y - ey, gb.z - ez, 0.0);
  glam[k] = iter == 0 ? 0.0 : gl;
}

__global__ __launch_bounds__(kBs) void WideBendAdjKernel(WideArgs a, int color, float inv_dt2,
                                                         BendMode m,
                                                         const float4* __restrict__ tape,
                                                         double4* __restrict__ gx,
                                                         double* __restrict__ gbcomp) {
#pragma clang fp contract(off)
  const InstDesc di = a.inst[blockIdx.y];
  if (color >= di.b_ncol) return;
  const int beg = a.b_coloff[di.b_cbeg + color], end = a.b_coloff[di.b_cbeg + color + 1];
  const int k = beg + blockIdx.x * kBs + threadIdx.x;
  if (k >= end) return;
  const int4 lid = a.bidx[k];
  const int id1 = di.vbeg + lid.x, id2 = di.vbeg + lid.y, id3 = di.vbeg + lid.z, id4 = di.vbeg + lid.w;
  const float4 t0 = tape[3 * k], t1 = tape[3 * k + 1], t2 = tape[3 * k + 2];
  const float3 fx1 = make_float3(t0.x, t0.y, t0.z), fx2 = make_float3(t0.w, t1.x, t1.y);
  const float3 fx3 = make_float3(t1.z, t1.w, t2.x), fx4 = make_float3(t2.y, t2.z, t2.w);
  const float w1 = a.xw[id1].w, w2 = a.xw[id2].w, w3 = a.xw[id3].w, w4 = a.xw[id4].w;
  const float2 par = a.bpar[k];
  float fdraw;
  {  // the forward's tests (BendSolve, FP32)
    const float3 p2 = Sub(fx2, fx1), p3 = Sub(fx3, fx1), p4 = Sub(fx4, fx1);
    const float3 c23 = Cross(p2, p3), c24 = Cross(p2, p4);
    const float3 n1 = Scale(c23, 1.0f / sqrtf(Dot(c23, c23)));
    const float3 n2 = Scale(c24, 1.0f / sqrtf(Dot(c24, c24)));
    fdraw = Dot(n1, n2);
    if (fdraw != fdraw) return;
    const float fd = fminf(fmaxf(fdraw, m.dlo), 1.0f);
    if (acosf(fd) < kEps) return;
    const float i23 = 1.0f / (sqrtf(Dot(c23, c23)) + kEps), i24 = 1.0f / (sqrtf(Dot(c24, c24)) + kEps);
    const float3 q3 = Scale(Add(Cross(p2, n2), Scale(Cross(n1, p2), fd)), i23);
    const float3 q4 = Scale(Add(Cross(p2, n1), Scale(Cross(n2, p2), fd)), i24);
    const float3 q2 = Sub(Scale(Add(Cross(p3, n2), Scale(Cross(n1, p3), fd)), -i23),
                          Scale(Add(Cross(p4, n1), Scale(Cross(n2, p4), fd)), i24));
    const float3 q1 = Scale(Add(Add(q2, q3), q4), -1.0f);
    const float fden = par.y * inv_dt2 +
                       (w1 * Dot(q1, q1) + w2 * Dot(q2, q2) + w3 * Dot(q3, q3) + w4 * Dot(q4, q4));
    if (fden < kEps) return;
  }
  // forward quantities in double
  const D3 x1 = d3(fx1), x2 = d3(fx2), x3 = d3(fx3), x4 = d3(fx4);
  const D3 p2 = Sb(x2, x1), p3 = Sb(x3, x1), p4 = Sb(x4, x1);
  const D3 c23 = Cr(p2, p3), c24 = Cr(p2, p4);
  const double l23r = sqrt(Dt(c23, c23)), l24r = sqrt(Dt(c24, c24));
  const D3 n1 = Sc(c23, 1.0 / l23r), n2 = Sc(c24, 1.0 / l24r);
  const double draw = Dt(n1, n2);
  const double d = fmin(fmax(draw, static_cast<double>(m.dlo)), 1.0);
  const double ang = acos(d);
  const double i23 = 1.0 / (l23r + kEps), i24 = 1.0 / (l24r + kEps);
  const D3 A3 = Ad(Cr(p2, n2), Sc(Cr(n1, p2), d));
  const D3 A4 = Ad(Cr(p2, n1), Sc(Cr(n2, p2), d));
  const D3 A2a = Ad(Cr(p3, n2), Sc(Cr(n1, p3), d));
  const D3 A2b = Ad(Cr(p4, n1), Sc(Cr(n2, p4), d));
  const D3 q3 = Sc(A3, i23), q4 = Sc(A4, i24);
  const D3 q2 = Sb(Sc(A2a, -i23), Sc(A2b, i24));
  const D3 q1 = Sc(Ad(Ad(q2, q3), q4), -1.0);
  const double denom = static_cast<double>(par.y) * inv_dt2 +
                       (w1 * Dt(q1, q1) + w2 * Dt(q2, q2) + w3 * Dt(q3, q3) + w4 * Dt(q4, q4));
  const double th = static_cast<double>(m.th0) + static_cast<double>(m.th1) * par.x;
  const double s = sqrt(fmax(1.0 - d * d, 0.0));
  const double lam = m.sgn * s * (ang - th) / denom;

  // reverse
  const D3 g1 = d3(gx[id1]), g2 = d3(gx[id2]), g3 = d3(gx[id3]), g4 = d3(gx[id4]);
  const double glam = w1 * Dt(q1, g1) + w2 * Dt(q2, g2) + w3 * Dt(q3, g3) + w4 * Dt(q4, g4);
  const double gden = -glam * lam / denom;
  gbcomp[k] += gden * inv_dt2;
  D3 gq1 = Ad(Sc(g1, w1 * lam), Sc(q1, 2.0 * gden * w1));
  D3 gq2 = Ad(Sc(g2, w2 * lam), Sc(q2, 2.0 * gden * w2));
  D3 gq3 = Ad(Sc(g3, w3 * lam), Sc(q3, 2.0 * gden * w3));
  D3 gq4 = Ad(Sc(g4, w4 * lam), Sc(q4, 2.0 * gden * w4));
  const double gP = glam * m.sgn / denom;
  double gd = s > 0.0 ? gP * (-d * (ang - th) / s - 1.0) : 0.0;
  // q1 = -(q2 + q3 + q4)
  Ac(gq2, Sc(gq1, -1.0));
  Ac(gq3, Sc(gq1, -1.0));
  Ac(gq4, Sc(gq1, -1.0));
  // q3 = A3 i23, q4 = A4 i24, q2 = -A2a i23 - A2b i24
  const D3 gA3 = Sc(gq3, i23), gA4 = Sc(gq4, i24);
  const D3 gA2a = Sc(gq2, -i23), gA2b = Sc(gq2, -i24);
  const double gi23 = Dt(gq3, A3) - Dt(gq2, A2a);
  const double gi24 = Dt(gq4, A4) - Dt(gq2, A2b);
  D3 gp2{0.0, 0.0, 0.0}, gp3 = gp2, gp4 = gp2, gn1 = gp2, gn2 = gp2;
  // A3 = p2 x n2 + d (n1 x p2)
  Ac(gp2, Cr(n2, gA3)); Ac(gn2, Cr(gA3, p2));
  gd += Dt(gA3, Cr(n1, p2));
  { const D3 t = Sc(gA3, d); Ac(gn1, Cr(p2, t)); Ac(gp2, Cr(t, n1)); }
  // A4 = p2 x n1 + d (n2 x p2)
  Ac(gp2, Cr(n1, gA4)); Ac(gn1, Cr(gA4, p2));
  gd += Dt(gA4, Cr(n2, p2));
  { const D3 t = Sc(gA4, d); Ac(gn2, Cr(p2, t)); Ac(gp2, Cr(t, n2)); }
  // A2a = p3 x n2 + d (n1 x p3)
  Ac(gp3, Cr(n2, gA2a)); Ac(gn2, Cr(gA2a, p3));
  gd += Dt(gA2a, Cr(n1, p3));
  { const D3 t = Sc(gA2a, d); Ac(gn1, Cr(p3, t)); Ac(gp3, Cr(t, n1)); }
  // A2b = p4 x n1 + d (n2 x p4)
  Ac(gp4, Cr(n1, gA2b)); Ac(gn1, Cr(gA2b, p4));
  gd += Dt(gA2b, Cr(n2, p4));
  { const D3 t = Sc(gA2b, d); Ac(gn2, Cr(p4, t)); Ac(gp4, Cr(t, n2)); }
  // i23 = 1 / (|c23| + eps): d|c23|/dc23 = n1
  const double gl23 = -gi23 * i23 * i23, gl24 = -gi24 * i24 * i24;
  // clamp (the forward's FP32 d decides the branch), then d = n1 . n2
  const double gdr = (fdraw < m.dlo || fdraw > 1.0f) ? 0.0 : gd;
  Ac(gn1, Sc(n2, gdr));
  Ac(gn2, Sc(n1, gdr));
  // n = c / |c|: gc = (gn - n (n . gn)) / |c| + g|c| n
  const D3 gc23 = Ad(Sc(Sb(gn1, Sc(n1, Dt(n1, gn1))), 1.0 / l23r), Sc(n1, gl23));
  const D3 gc24 = Ad(Sc(Sb(gn2, Sc(n2, Dt(n2, gn2))), 1.0 / l24r), Sc(n2, gl24));
  // c23 = p2 x p3, c24 = p2 x p4
  Ac(gp2, Cr(p3, gc23)); Ac(gp3, Cr(gc23, p2));
  Ac(gp2, Cr(p4, gc24)); Ac(gp4, Cr(gc24, p2));
  // p_k = x_k - x1, plus the identity part of x' = x + w lam q
  gx[id1] = d4(Sb(g1, Ad(Ad(gp2, gp3), gp4)));
  gx[id2] = d4(Ad(g2, gp2));
  gx[id3] = d4(Ad(g3, gp3));
  gx[id4] = d4(Ad(g4, gp4));
}

// LRA (reverse): one target vertex of any instance per thread.  y = x - u + u R / |u|
// (u = x - s, R = rest x stretchiness; active when |u| > R): dy/dx = (R / |u|)(I - u^ u^T),
// dy/ds = I - dy/dx.  The forward replay repeats LraVertex's FP32 arithmetic, so the
// activity decisions and the replayed positions are the forward's.
__global__ __launch_bounds__(kBs) void WideLraAdjKernel(const InstDesc* __restrict__ inst,
                                                        const float4* __restrict__ xpre,
                                                        const int32_t* __restrict__ slot_vert,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ slot,
                                                        const float* __restrict__ rest,
                                                        float stretchiness,
                                                        float4* __restrict__ xs,
                                                        double4* __restrict__ gslot,
                                                        double4* __restrict__ gx) {
#pragma clang fp contract(off)
  const InstDesc d = inst[blockIdx.y];
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (d.nslots == 0 || i >= d.nv) return;
  const int kb = ptr[d.lra_vbeg + i], ke = ptr[d.lra_vbeg + i + 1];
  if (kb == ke) return;
  const float4* X = xpre + d.vbeg;
  const int32_t* sv = slot_vert + d.slot_beg;
  float4 xt = X[i];
  for (int k = kb; k < ke; ++k) {  // forward replay (LraVertex)
    xs[k] = xt;
    const float target = rest[k] * stretchiness;
    if (xt.w < kEps && target > 0.0f) continue;
    const float4 sp = X[sv[slot[k]]];
    const float dx = xt.x - sp.x, dy = xt.y - sp.y, dz = xt.z - sp.z;
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    if (dist > target) {
      const float f = target / dist;
      xt.x += -dx + dx * f; xt.y += -dy + dy * f; xt.z += -dz + dz * f;
    }
  }
  double4 g = gx[d.vbeg + i];
  for (int k = ke - 1; k >= kb; --k) {
    const float4 a = xs[k];
    const float target = rest[k] * stretchiness;
    double4 gs = make_double4(0.0, 0.0, 0.0, 0.0);
    if (!(a.w < kEps && target > 0.0f)) {
      const float4 sp = X[sv[slot[k]]];
      const float fx = a.x - sp.x, fy = a.y - sp.y, fz = a.z - sp.z;
      if (sqrtf(fx * fx + fy * fy + fz * fz) > target) {
        const double dx = static_cast<double>(a.x) - sp.x, dy = static_cast<double>(a.y) - sp.y,
                     dz = static_cast<double>(a.z) - sp.z;
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        const double ux = dx / dist, uy = dy / dist, uz = dz / dist;
        const double f = static_cast<double>(target) / dist;
        const double ug = ux * g.x + uy * g.y + uz * g.z;
        const double nx = f * (g.x - ux * ug), ny = f * (g.y - uy * ug), nz = f * (g.z - uz * ug);
        gs = make_double4(g.x - nx, g.y - ny, g.z - nz, 0.0);
        g = make_double4(nx, ny, nz, 0.0);
      }
    }
    gslot[k] = gs;
  }
  gx[d.vbeg + i] = g;
}

// LRA (reverse), slots: one workgroup per global slot sums its constraints' adjoints in a
// fixed order (strided partial sums, then a tree in LDS) into the slot vertex's gx
__global__ __launch_bounds__(kBs) void LraSlotAdjKernel(const InstDesc* __restrict__ inst,
                                                        const int32_t* __restrict__ slot_vert,
                                                        const int32_t* __restrict__ slot_inst,
                                                        const int32_t* __restrict__ slot_ptr,
                                                        const int32_t* __restrict__ slot_cons,
                                                        const double4* __restrict__ gslot,
                                                        double4* __restrict__ gx) {
  __shared__ D3 part[kBs];
  const int gs = blockIdx.x;
  D3 acc{0.0, 0.0, 0.0};
  for (int j = slot_ptr[gs] + threadIdx.x; j < slot_ptr[gs + 1]; j += kBs)
    Ac(acc, d3(gslot[slot_cons[j]]));
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBs / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) Ac(part[threadIdx.x], part[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const InstDesc d = inst[slot_inst[gs]];
    double4* p = gx + d.vbeg + slot_vert[gs];
    *p = d4(Ad(d3(*p), part[0]));
  }
}

__global__ __launch_bounds__(kBs) void ZeroD4Kernel(double4* __restrict__ p, int n) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i < n) p[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}
__global__ __launch_bounds__(kBs) void ZeroD1Kernel(double* __restrict__ p, int n) {
  const int i = blockIdx.x * kBs + threadIdx.x;
  if (i < n) p[i] = 0.0;
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
hipError_t LaunchWideLraAdj(hipStream_t s, const InstDesc* inst, int n_inst, int max_nv,
                            int total_slots, const float4* xpre, const int32_t* slot_vert,
                            const int32_t* ptr, const int32_t* slot, const float* rest,
                            float stretchiness, float4* xs, double4* gslot,
                            const int32_t* slot_inst, const int32_t* slot_ptr,
                            const int32_t* slot_cons, double4* gx) {
  if (n_inst <= 0 || max_nv <= 0 || total_slots <= 0) return hipSuccess;
  WideLraAdjKernel<<<dim3(Blocks(max_nv), n_inst), kBs, 0, s>>>(inst, xpre, slot_vert, ptr, slot,
                                                                rest, stretchiness, xs, gslot, gx);
  LraSlotAdjKernel<<<total_slots, kBs, 0, s>>>(inst, slot_vert, slot_inst, slot_ptr, slot_cons,
                                               gslot, gx);
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
hipError_t LaunchZeroD4(hipStream_t s, double4* p, int n) {
  if (n <= 0) return hipSuccess;
  ZeroD4Kernel<<<Blocks(n), kBs, 0, s>>>(p, n);
  return hipGetLastError();
}
hipError_t LaunchZeroD1(hipStream_t s, double* p, int n) {
  if (n <= 0) return hipSuccess;
  ZeroD1Kernel<<<Blocks(n), kBs, 0, s>>>(p, n);
  return hipGetLastError();
}
hipError_t LaunchFinalizeAdj(hipStream_t s, double4* gx0, const double4* gv, double4* gx, int n,
                             double f) {
  if (n <= 0) return hipSuccess;
  FinalizeAdjKernel<<<Blocks(n), kBs, 0, s>>>(gx0, gv, gx, n, f);
  return hipGetLastError();
}
hipError_t LaunchPredictAdj(hipStream_t s, double4* gx0, double4* gv, const double4* gx, int n,
                            double dt) {
  if (n <= 0) return hipSuccess;
  PredictAdjKernel<<<Blocks(n), kBs, 0, s>>>(gx0, gv, gx, n, dt);
  return hipGetLastError();
}
hipError_t LaunchGroundAdj(hipStream_t s, const float4* xpre, const float4* x0s, double4* gx,
                           double4* gx0, int n, float ground_y, float mu) {
  if (n <= 0) return hipSuccess;
  GroundAdjKernel<<<Blocks(n), kBs, 0, s>>>(xpre, x0s, gx, gx0, n, ground_y, mu);
  return hipGetLastError();
}
hipError_t LaunchWideStretchAdj(hipStream_t s, const WideArgs& a, int n_inst, int color,
                                int max_cnt, float inv_dt2, int iter, const float4* tape,
                                double4* gx, double* glam, double* gcomp) {
  if (max_cnt <= 0 || n_inst <= 0) return hipSuccess;
  WideStretchAdjKernel<<<dim3(Blocks(max_cnt), n_inst), kBs, 0, s>>>(a, color, inv_dt2, iter, tape,
                                                                      gx, glam, gcomp);
  return hipGetLastError();
}
hipError_t LaunchWideBendAdj(hipStream_t s, const WideArgs& a, int n_inst, int color, int max_cnt,
                             float inv_dt2, const BendMode& m, const float4* tape, double4* gx,
                             double* gbcomp) {
  if (max_cnt <= 0 || n_inst <= 0) return hipSuccess;
  WideBendAdjKernel<<<dim3(Blocks(max_cnt), n_inst), kBs, 0, s>>>(a, color, inv_dt2, m, tape, gx,
                                                                   gbcomp);
  return hipGetLastError();
}

}  // namespace xrt
