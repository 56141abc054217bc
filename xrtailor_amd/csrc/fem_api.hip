// Quality-mode FEM constraints: set-up (FEM::StrainSolver and FEM::IsometricBendingSolver,
// xrtailor/src/physics/dynamics/FEMSolver.cu:9-86, with the rest data of InitStrain_Kernel and
// InitIsometricBending_Kernel, FEMHelper.cu:10-94, 221-301) and the per-iteration drivers
// (ClothSolverBase::SimulateQuality, xrtailor/src/pipeline/base/ClothSolverBase.cpp:314-323).
//
// The reference builds its rest data on the GPU in FP32; here they are computed once on the
// host in FP64 and stored in FP32.  Faces are colored like every other constraint set (the
// context's coloring mode and seed), then stored color-sorted per instance; the isometric
// bending data share the bend elements' color order, so the bend colors and the wide bend
// block maps serve both constraint models (the wide strain launch, two faces per thread, has
// block maps of its own).
#include <algorithm>
#include <cmath>

#include "context.h"
#include "fem.h"
#include "topology.h"

namespace xrt {
namespace {

struct V3 {
  double x, y, z;
};
V3 At(const float* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 Cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 Unit(V3 a) {
  const double l = std::sqrt(Dot(a, a));
  return {a.x / l, a.y / l, a.z / l};
}

// ComputeStrainMeta (FEMHelper.cu:10-40): the vertices in the face's own 2D frame,
// q_i = (p_i . n x (p1 - p0), p_i . (p1 - p0)) with unit axes, rest matrix P whose column 0
// holds the first frame coordinate of q0 - q2 and q1 - q2 and column 1 the second, its
// inverse in glm order; the identity for a degenerate face
StrainC StrainRest(const float* pos, int i0, int i1, int i2) {
  const V3 p0 = At(pos, i0), p1 = At(pos, i1), p2 = At(pos, i2);
  const V3 n = Cross(Sub(p1, p0), Sub(p2, p0));
  const V3 ax = Unit(Sub(p1, p0));
  const V3 ay = Unit(Cross(n, ax));
  const double u0 = Dot(p0, ay), v0 = Dot(p0, ax), u1 = Dot(p1, ay), v1 = Dot(p1, ax);
  const double u2 = Dot(p2, ay), v2 = Dot(p2, ax);
  const double c0r0 = u0 - u2, c0r1 = u1 - u2;  // column 0
  const double c1r0 = v0 - v2, c1r1 = v1 - v2;  // column 1
  const double det = c0r0 * c1r1 - c0r1 * c1r0;
  StrainC c{};
  c.i0 = i0;
  c.i1 = i1;
  c.i2 = i2;
  c.area = static_cast<float>(0.5 * std::sqrt(Dot(n, n)));
  if (std::fabs(det) > 1e-6) {
    c.inv[0] = static_cast<float>(c1r1 / det);
    c.inv[1] = static_cast<float>(-c0r1 / det);
    c.inv[2] = static_cast<float>(-c1r0 / det);
    c.inv[3] = static_cast<float>(c0r0 / det);
  } else {
    c.inv[0] = 1.0f;
    c.inv[1] = 0.0f;
    c.inv[2] = 0.0f;
    c.inv[3] = 1.0f;
  }
  return c;
}

double Cot(V3 a, V3 b) {
  const V3 c = Cross(a, b);
  return Dot(a, b) / std::sqrt(Dot(c, c));
}

// ComputeHessianEnergy (FEMHelper.cu:221-245): Q = coeff K K^T over the hinge (b0, b1 the
// shared edge, b2 and b3 the opposite vertices), stored as K and coeff.  A degenerate hinge
// (non-finite weights) is stored inert (K = 0).
void IsoRest(const float* pos, const int32_t* b, float4* K, float* coeff) {
  const V3 p0 = At(pos, b[0]), p1 = At(pos, b[1]), p2 = At(pos, b[2]), p3 = At(pos, b[3]);
  const V3 e0 = Sub(p1, p0), e1 = Sub(p2, p0), e2 = Sub(p3, p0), e3 = Sub(p2, p1),
           e4 = Sub(p3, p1);
  const V3 me0 = {-e0.x, -e0.y, -e0.z};
  const double c01 = Cot(e0, e1), c02 = Cot(e0, e2), c03 = Cot(me0, e3), c04 = Cot(me0, e4);
  const V3 n0 = Cross(e0, e1), n1 = Cross(Sub(p0, p1), Sub(p3, p1));
  const double a0 = 0.5 * std::sqrt(Dot(n0, n0)), a1 = 0.5 * std::sqrt(Dot(n1, n1));
  const double k[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
  const double cf = 3.0 / (a0 + a1);
  bool ok = std::isfinite(cf);
  for (double v : k) ok = ok && std::isfinite(v);
  if (!ok) {
    *K = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    *coeff = 0.0f;
    return;
  }
  *K = make_float4(static_cast<float>(k[0]), static_cast<float>(k[1]), static_cast<float>(k[2]),
                   static_cast<float>(k[3]));
  *coeff = static_cast<float>(cf);
}

// the orthotropic elasticity entries (FEMHelper.cu:103-111) and the bending alpha-tilde
// 1 / (stiffness dt^2) (FEMHelper.cu:360) in FP32
FemConsts MakeFemConsts(const xrt_params& p, float dt) {
  FemConsts k{};
  const float den = 1.0f - p.fem_xy_poisson_ratio * p.fem_yx_poisson_ratio;
  k.c00 = p.fem_xx_stiffness / den;
  k.c01 = p.fem_xx_stiffness * p.fem_yx_poisson_ratio / den;
  k.c10 = p.fem_yy_stiffness * p.fem_xy_poisson_ratio / den;
  k.c11 = p.fem_yy_stiffness / den;
  k.c22 = p.fem_xy_stiffness;
  k.iso_at = 1.0f / (p.fem_bending_stiffness * dt * dt);
  return k;
}

float SubstepDt(const xrt_params& p) { return p.frame_time / static_cast<float>(p.num_substeps); }

FemArgs MakeFemArgs(xrt_ctx* c) {
  FemData& F = c->fem;
  FemArgs a{};
  a.inst = c->desc.ptr;
  a.fcol = F.fcol.ptr;
  a.f_coloff = F.f_coloff.ptr;
  a.sc = F.sc.ptr;
  a.b_coloff = c->b_coloff.ptr;
  a.bidx = c->bidx.ptr;
  a.iso_k = F.iso_k.ptr;
  a.iso_c = F.iso_c.ptr;
  a.iso_lam = F.iso_lam.ptr;
  a.x = c->x.ptr;
  return a;
}

int CheckFem(xrt_ctx* c, int32_t inst) {
  XRT_REQUIRE(c != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(inst >= 0 && inst < static_cast<int32_t>(c->inst.size()), XRT_ERR_INVALID_ARG,
              "instance index out of range");
  XRT_REQUIRE(c->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(FemOn(c) && c->fem.ready, XRT_ERR_STATE, "constraint_model is not FEM");
  return XRT_OK;
}
}  // namespace

bool FemOn(const xrt_ctx* ctx) { return ctx->params.constraint_model == XRT_CONSTRAINTS_FEM; }

FemConsts FemConstsOf(const xrt_params& p) { return MakeFemConsts(p, SubstepDt(p)); }

int FemPrepare(xrt_ctx* ctx) {
  FemData& F = ctx->fem;
  if (!FemOn(ctx) || F.ready) return XRT_OK;
  const int ni = static_cast<int>(ctx->inst.size());
  F.inst.assign(ni, FemData::Inst{});
  std::vector<StrainC> sc;
  std::vector<int32_t> coloff;
  std::vector<int2> fcol(ni);
  int b_total = 0;
  for (int i = 0; i < ni; ++i) {
    const HostInstance& h = ctx->inst[i];
    FemData::Inst& fi = F.inst[i];
    fi.color = ColorConstraints(h.tris.data(), 3, h.nt, h.nv, F.seed + static_cast<uint64_t>(i),
                                ctx->params.coloring, &fi.ncol);
    ColorSort(fi.color, fi.ncol, &fi.order, &fi.coloff);
    fi.fbeg = static_cast<int>(sc.size());
    fcol[i] = make_int2(static_cast<int>(coloff.size()), fi.ncol);
    for (int c = 0; c <= fi.ncol; ++c) coloff.push_back(fi.fbeg + fi.coloff[c]);
    for (int p = 0; p < h.nt; ++p) {
      const int f = fi.order[p];
      sc.push_back(StrainRest(h.rest_pos.data(), h.tris[3 * f], h.tris[3 * f + 1], h.tris[3 * f + 2]));
    }
    b_total = std::max(b_total, h.bbeg + h.nb());
  }
  std::vector<float4> iso_k(b_total);
  std::vector<float> iso_c(b_total);
  for (const HostInstance& h : ctx->inst)
    for (int p = 0; p < h.nb(); ++p)
      IsoRest(h.rest_pos.data(), &h.bends[4 * h.b_order[p]], &iso_k[h.bbeg + p], &iso_c[h.bbeg + p]);
  // per-color block maps of the wide strain launches (n + 1 prefix sums, then the instance
  // of every block; kFemPerBlock faces per block; only blocks with faces are launched)
  int max_col = 0;
  for (const auto& fi : F.inst) max_col = std::max(max_col, fi.ncol);
  std::vector<int32_t> maps;
  F.f_map.assign(max_col, 0);
  F.f_blocks.assign(max_col, 0);
  for (int col = 0; col < max_col; ++col) {
    std::vector<int32_t> o(ni + 1, 0);
    for (int i = 0; i < ni; ++i) {
      const FemData::Inst& fi = F.inst[i];
      const int cnt = col < fi.ncol ? fi.coloff[col + 1] - fi.coloff[col] : 0;
      o[i + 1] = o[i] + (cnt + kFemPerBlock - 1) / kFemPerBlock;
    }
    F.f_map[col] = maps.size();
    F.f_blocks[col] = o[ni];
    maps.insert(maps.end(), o.begin(), o.end());
    for (int i = 0; i < ni; ++i) maps.insert(maps.end(), static_cast<size_t>(o[i + 1] - o[i]), i);
  }
  hipStream_t s = ctx->stream;
  XRT_CHECK_HIP(F.sc.Upload(sc, s));
  XRT_CHECK_HIP(F.f_coloff.Upload(coloff, s));
  XRT_CHECK_HIP(F.fcol.Upload(fcol, s));
  XRT_CHECK_HIP(F.iso_k.Upload(iso_k, s));
  XRT_CHECK_HIP(F.iso_c.Upload(iso_c, s));
  XRT_CHECK_HIP(F.iso_lam.Alloc(std::max(b_total, 1)));
  XRT_CHECK_HIP(hipMemsetAsync(F.iso_lam.ptr, 0, sizeof(float) * std::max(b_total, 1), s));
  XRT_CHECK_HIP(F.maps.Upload(maps, s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  F.ready = true;
  return XRT_OK;
}

int FemIter(xrt_ctx* c, int inst, int it) {
  const FemData& F = c->fem;
  XRT_REQUIRE(F.ready, XRT_ERR_STATE, "FEM constraints are not set up");
  const FemConsts k = MakeFemConsts(c->params, SubstepDt(c->params));
  const bool bend = c->params.solve_bending != 0;
  hipStream_t s = c->stream;
  if (inst >= 0) {
    const HostInstance& h = c->inst[inst];
    const FemData::Inst& fi = F.inst[inst];
    float4* x = c->x.ptr + h.desc.vbeg;
    for (int col = 0; col < fi.ncol; ++col)
      XRT_CHECK_HIP(LaunchStrain(s, F.sc.ptr + fi.fbeg + fi.coloff[col], x,
                                 fi.coloff[col + 1] - fi.coloff[col], k));
    if (bend)
      for (int col = 0; col < h.b_ncol; ++col) {
        const int b = h.bbeg + h.b_coloff[col];
        XRT_CHECK_HIP(LaunchIsoBend(s, c->bidx.ptr + b, F.iso_k.ptr + b, F.iso_c.ptr + b,
                                    F.iso_lam.ptr + b, x, h.b_coloff[col + 1] - h.b_coloff[col],
                                    it, k));
      }
    return XRT_OK;
  }
  const int ni = static_cast<int>(c->inst.size());
  const FemArgs a = MakeFemArgs(c);
  for (size_t col = 0; col < F.f_blocks.size(); ++col)
    XRT_CHECK_HIP(LaunchWideStrain(s, a, ni, static_cast<int>(col), F.maps.ptr + F.f_map[col],
                                   F.f_blocks[col], k));
  if (bend) {
    const WideConfig& W = c->wide;
    for (size_t col = 0; col < W.b_blocks.size(); ++col)
      XRT_CHECK_HIP(LaunchWideIsoBend(s, a, ni, static_cast<int>(col), W.maps.ptr + W.b_map[col],
                                      W.b_blocks[col], it, k));
  }
  return XRT_OK;
}

}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_strain_solve(xrt_ctx* ctx, int32_t inst, int32_t color) {
  int rc = CheckFem(ctx, inst);
  if (rc != XRT_OK) return rc;
  const FemData::Inst& fi = ctx->fem.inst[inst];
  XRT_REQUIRE(color >= 0 && color < fi.ncol, XRT_ERR_INVALID_ARG, "color out of range");
  const FemConsts k = MakeFemConsts(ctx->params, SubstepDt(ctx->params));
  XRT_CHECK_HIP(LaunchStrain(ctx->stream, ctx->fem.sc.ptr + fi.fbeg + fi.coloff[color],
                             ctx->x.ptr + ctx->inst[inst].desc.vbeg,
                             fi.coloff[color + 1] - fi.coloff[color], k));
  return XRT_OK;
}

int xrt_isometric_bending_solve(xrt_ctx* ctx, int32_t inst, int32_t color, int32_t iter, float dt) {
  int rc = CheckFem(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  XRT_REQUIRE(color >= 0 && color < h.b_ncol, XRT_ERR_INVALID_ARG, "color out of range");
  XRT_REQUIRE(dt > 0.0f && std::isfinite(dt), XRT_ERR_INVALID_ARG, "dt must be finite and > 0");
  XRT_REQUIRE(iter >= 0, XRT_ERR_INVALID_ARG, "iter must be >= 0");
  const FemConsts k = MakeFemConsts(ctx->params, dt);
  const FemData& F = ctx->fem;
  const int b = h.bbeg + h.b_coloff[color];
  XRT_CHECK_HIP(LaunchIsoBend(ctx->stream, ctx->bidx.ptr + b, F.iso_k.ptr + b, F.iso_c.ptr + b,
                              F.iso_lam.ptr + b, ctx->x.ptr + h.desc.vbeg,
                              h.b_coloff[color + 1] - h.b_coloff[color], iter, k));
  return XRT_OK;
}

int xrt_get_fem(xrt_ctx* ctx, int32_t inst, int32_t* face_colors, int32_t* n_face_colors,
                float* area, float* inv, float* iso_k, float* iso_coeff) {
  int rc = CheckFem(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  const FemData::Inst& fi = ctx->fem.inst[inst];
  if (face_colors) std::copy(fi.color.begin(), fi.color.end(), face_colors);
  if (n_face_colors) *n_face_colors = fi.ncol;
  if (area || inv) {
    std::vector<StrainC> sc(h.nt);
    if (h.nt > 0)
      XRT_CHECK_HIP(hipMemcpyAsync(sc.data(), ctx->fem.sc.ptr + fi.fbeg, sizeof(StrainC) * h.nt,
                                   hipMemcpyDeviceToHost, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < h.nt; ++p) {
      const int f = fi.order[p];
      if (area) area[f] = sc[p].area;
      if (inv) std::copy(sc[p].inv, sc[p].inv + 4, inv + 4 * f);
    }
  }
  if (iso_k || iso_coeff) {
    const int nb = h.nb();
    std::vector<float4> k(nb);
    std::vector<float> cf(nb);
    if (nb > 0) {
      XRT_CHECK_HIP(hipMemcpyAsync(k.data(), ctx->fem.iso_k.ptr + h.bbeg, sizeof(float4) * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
      XRT_CHECK_HIP(hipMemcpyAsync(cf.data(), ctx->fem.iso_c.ptr + h.bbeg, sizeof(float) * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
    }
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < nb; ++p) {
      const int b = h.b_order[p];
      if (iso_k) {
        iso_k[4 * b] = k[p].x;
        iso_k[4 * b + 1] = k[p].y;
        iso_k[4 * b + 2] = k[p].z;
        iso_k[4 * b + 3] = k[p].w;
      }
      if (iso_coeff) iso_coeff[b] = cf[p];
    }
  }
  return XRT_OK;
}

int xrt_time_fem(xrt_ctx* ctx, int32_t n, float ms[2]) {
  XRT_REQUIRE(ctx && ms && n > 0, XRT_ERR_INVALID_ARG, "null argument or n <= 0");
  XRT_REQUIRE(ctx->started && FemOn(ctx) && ctx->fem.ready, XRT_ERR_STATE,
              "constraint_model is not FEM");
  const FemData& F = ctx->fem;
  const WideConfig& W = ctx->wide;
  const FemConsts k = MakeFemConsts(ctx->params, SubstepDt(ctx->params));
  const FemArgs a = MakeFemArgs(ctx);
  const int ni = static_cast<int>(ctx->inst.size());
  hipStream_t s = ctx->stream;
  // every launch bracketed by its own event pair (the gaps between launches excluded)
  double sum[2] = {0.0, 0.0};
  int cnt[2] = {0, 0};
  for (int kind = 0; kind < 2; ++kind)
    for (int r = 0; r < n; ++r) {
      const int ncol = kind == 0 ? static_cast<int>(F.f_blocks.size())
                                 : static_cast<int>(W.b_blocks.size());
      for (int col = 0; col < ncol; ++col) {
        XRT_CHECK_HIP(hipEventRecord(ctx->ev0, s));
        if (kind == 0)
          XRT_CHECK_HIP(LaunchWideStrain(s, a, ni, col, F.maps.ptr + F.f_map[col],
                                         F.f_blocks[col], k));
        else
          XRT_CHECK_HIP(LaunchWideIsoBend(s, a, ni, col, W.maps.ptr + W.b_map[col],
                                          W.b_blocks[col], r, k));
        XRT_CHECK_HIP(hipEventRecord(ctx->ev1, s));
        XRT_CHECK_HIP(hipEventSynchronize(ctx->ev1));
        float t = 0.0f;
        XRT_CHECK_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        sum[kind] += t;
        cnt[kind]++;
      }
    }
  ms[0] = static_cast<float>(sum[0] / std::max(cnt[0], 1));
  ms[1] = static_cast<float>(sum[1] / std::max(cnt[1], 1));
  return XRT_OK;
}

}  // extern "C"
