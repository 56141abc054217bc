// Differentiable rollout: the wide forward driver (every instance per launch, optional
// tape), checkpointing with recompute, the reverse driver, and the C ABI.
//
// Schedule of xrt_adjoint_backward over F frames with checkpoints every K frames:
//   for each segment (last first): restore its checkpoint, run its frames forward keeping
//   each frame's start state; then for each frame of the segment (last first): restore the
//   frame start, run the frame forward with the tape on, reverse its substeps (last first).
// The forward is deterministic (no float atomics; contact deltas in fixed point), so every
// recompute reproduces the recorded trajectory and its contact records bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "adjoint.h"
#include "contact.h"
#include "context.h"
#include "dynamics.h"

namespace xrt {
namespace {

constexpr size_t kMaxTapeBytes = size_t(96) << 30;  // tape and records of every frame slot

WideArgs MakeWideArgs(xrt_ctx* c) {
  WideArgs a{};
  a.inst = c->desc.ptr;
  a.s_coloff = c->s_coloff.ptr;
  a.b_coloff = c->b_coloff.ptr;
  a.sc = c->sc.ptr;
  a.slam = c->slam.ptr;
  a.bidx = c->bidx.ptr;
  a.bpar = c->bpar.ptr;
  a.x = c->x.ptr;
  a.xw = c->x0.ptr;
  return a;
}

bool ObsContact(const xrt_ctx* c, size_t i) {
  return c->contact[i] && c->contact[i]->obs_nf > 0 && c->params.enable_obstacle_contact;
}

// validates the scene for the adjoint and sizes every buffer except the checkpoints; the
// tape and the contact records get `want_slots` frame slots when they fit in 96 GiB and in
// 60 % of the free device memory, else one
int AdjSetup(xrt_ctx* c, int want_slots = 1) {
  XRT_REQUIRE(c->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(!c->params.enable_particle_collision, XRT_ERR_UNSUPPORTED,
              "the adjoint does not support particle self-collision");
  XRT_REQUIRE(!BindingsOn(c), XRT_ERR_UNSUPPORTED, "the adjoint does not support bindings");
  XRT_REQUIRE(!FemOn(c), XRT_ERR_UNSUPPORTED, "the adjoint does not support the FEM constraints");
  XRT_REQUIRE(!(c->params.enable_obstacle_contact && c->params.contact_friction > 0.0f),
              XRT_ERR_UNSUPPORTED, "the adjoint does not support contact friction");
  const xrt_params& p = c->params;
  const int n_inst = static_cast<int>(c->inst.size());
  XRT_REQUIRE(n_inst > 0 && n_inst <= 65535, XRT_ERR_INVALID_ARG, "1..65535 instances");
  if (!c->adj) c->adj = std::make_unique<AdjointData>();
  AdjointData& A = *c->adj;
  const WideConfig& W = c->wide;
  A.n_total = W.n_total;
  A.e_total = W.e_total;
  A.b_total = W.b_total;
  A.s_max = W.s_max;
  A.b_max = W.b_max;
  A.contacts = false;
  for (int i = 0; i < n_inst; ++i) {
    if (c->contact[i]) {
      XRT_REQUIRE(c->contact[i]->frame_substep == 0, XRT_ERR_STATE,
                  "the adjoint starts at a frame boundary");
      XRT_REQUIRE(!c->contact[i]->obstacle_moving, XRT_ERR_UNSUPPORTED,
                  "the adjoint needs the obstacle at rest (xrt_update_obstacle since the last frame end)");
    }
    if (ObsContact(c, i)) A.contacts = true;
  }
  const size_t N = static_cast<size_t>(A.n_total);
  const bool bend = p.solve_bending != 0, ground = p.enable_ground != 0;
  A.lra = c->total_slots > 0;
  if (A.lra) {  // slot -> constraints (transpose of the per-target CSR), in constraint order
    std::vector<int32_t> cnt(c->total_slots + 1, 0), sinst(c->total_slots), cons;
    size_t total_k = 0;
    for (int i = 0; i < n_inst; ++i) total_k += c->inst[i].lra_slot.size();
    for (int i = 0; i < n_inst; ++i) {
      const HostInstance& h = c->inst[i];
      for (int j = 0; j < h.desc.nslots; ++j) sinst[h.desc.slot_beg + j] = i;
      for (int32_t q : h.lra_slot) cnt[h.desc.slot_beg + q + 1]++;
    }
    for (int g = 0; g < c->total_slots; ++g) cnt[g + 1] += cnt[g];
    cons.assign(std::max<size_t>(total_k, 1), 0);
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n_inst; ++i) {
      const HostInstance& h = c->inst[i];
      for (size_t q = 0; q < h.lra_slot.size(); ++q)
        cons[fill[h.desc.slot_beg + h.lra_slot[q]]++] = h.kbeg + static_cast<int32_t>(q);
    }
    XRT_CHECK_HIP(A.slot_inst.Upload(sinst, c->stream));
    XRT_CHECK_HIP(A.slot_ptr.Upload(cnt, c->stream));
    XRT_CHECK_HIP(A.slot_cons.Upload(cons, c->stream));
    if (A.lra_xs.n < cons.size()) {
      XRT_CHECK_HIP(A.lra_xs.Alloc(cons.size()));
      XRT_CHECK_HIP(A.lra_gslot.Alloc(cons.size()));
    }
  }
  A.off_s = 0;
  A.off_b = A.off_s + 2 * static_cast<size_t>(A.e_total);
  A.off_l = A.off_b + (bend ? 3 * static_cast<size_t>(A.b_total) : 0);
  A.off_g = A.off_l + (A.lra ? N : 0);
  A.off_c = A.off_g + (ground ? N : 0);
  A.off_e = A.off_c + (A.contacts ? N : 0);
  A.rep = p.enable_self_contact != 0;
  A.off_r = A.off_e + (A.contacts ? N : 0);
  A.it_stride = A.off_r + (A.rep ? N : 0);
  A.sub_stride = N + static_cast<size_t>(p.num_iterations) * A.it_stride;
  const size_t frame_n = static_cast<size_t>(p.num_substeps) * A.sub_stride;
  A.frequency = std::max(1, p.num_substeps * p.num_iterations / p.num_collision_passes);
  A.n_passes = (p.num_substeps * p.num_iterations + A.frequency - 1) / A.frequency;
  size_t rec_bytes = 0;  // contact records of one frame, every instance
  for (int i = 0; i < n_inst; ++i) {
    if (!ObsContact(c, i)) continue;
    const HostInstance& h = c->inst[i];
    rec_bytes += static_cast<size_t>(A.n_passes) *
                 (static_cast<size_t>(h.nv) * 40 + 16 * static_cast<size_t>(c->contact[i]->obs_nv) +
                  sizeof(EeRec) * std::max(h.ne(), 1));
  }
  // repulsion responses of one frame: at most 8 per node and pass (at least 2^20, at most
  // 2^27 entries of 16 B); more is reported as XRT_ERR_CAPACITY by the reverse pass
  A.rep_cap = 0;
  if (A.rep)
    A.rep_cap = std::min<unsigned long long>(
        1ull << 27, std::max<unsigned long long>(
                        1ull << 20, 8ull * N * static_cast<unsigned long long>(p.num_substeps) *
                                        static_cast<unsigned long long>(std::max(1, p.num_iterations))));
  const size_t frame_bytes = frame_n * sizeof(float4) + rec_bytes + A.rep_cap * sizeof(int4);
  XRT_REQUIRE(frame_bytes <= kMaxTapeBytes, XRT_ERR_CAPACITY,
              "one frame of adjoint tape exceeds 96 GiB");
  size_t free_b = 0, total_b = 0;
  XRT_CHECK_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t held = A.tape.n * sizeof(float4);  // reusable
  const size_t budget = std::min(kMaxTapeBytes, (free_b + held) / 10 * 6);
  A.slots = (want_slots > 1 && static_cast<size_t>(want_slots) * frame_bytes <= budget) ? want_slots : 1;
  A.slot = 0;
  const size_t tape_n = static_cast<size_t>(A.slots) * frame_n;
  if (A.tape.n != tape_n) XRT_CHECK_HIP(A.tape.Alloc(tape_n));
  A.rec.clear();
  A.rec.resize(n_inst);
  for (int i = 0; i < n_inst; ++i) {
    if (!ObsContact(c, i)) continue;
    const HostInstance& h = c->inst[i];
    const ContactData& cd = *c->contact[i];
    auto r = std::make_unique<AdjointData::Records>();
    const size_t np = static_cast<size_t>(A.n_passes) * A.slots;
    XRT_CHECK_HIP(r->vf_face.Alloc(np * h.nv));
    XRT_CHECK_HIP(r->rt_face.Alloc(np * h.nv));
    XRT_CHECK_HIP(r->vf_q.Alloc(np * h.nv));
    XRT_CHECK_HIP(r->rt_q.Alloc(np * h.nv));
    XRT_CHECK_HIP(r->pred.Alloc(np * cd.obs_nv));
    XRT_CHECK_HIP(r->ee.Alloc(np * sizeof(EeRec) * std::max(h.ne(), 1)));
    // slots of passes that have not run hold "no record" (face -1) and zeros
    XRT_CHECK_HIP(hipMemsetAsync(r->vf_face.ptr, 0xff, 4 * r->vf_face.n, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(r->rt_face.ptr, 0xff, 4 * r->rt_face.n, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(r->vf_q.ptr, 0, 16 * r->vf_q.n, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(r->rt_q.ptr, 0, 16 * r->rt_q.n, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(r->pred.ptr, 0, 16 * r->pred.n, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(r->ee.ptr, 0, r->ee.n, c->stream));
    {  // vertex -> edge CSR of the contact edges (ascending edge id per vertex)
      const int ne = h.ne();
      std::vector<int32_t> ptr(h.nv + 1, 0), adj(2 * static_cast<size_t>(ne));
      for (int e = 0; e < 2 * ne; ++e) ptr[h.edges[e] + 1]++;
      for (int v = 0; v < h.nv; ++v) ptr[v + 1] += ptr[v];
      std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
      for (int e = 0; e < ne; ++e) {
        adj[fill[h.edges[2 * e]]++] = e;
        adj[fill[h.edges[2 * e + 1]]++] = e;
      }
      XRT_CHECK_HIP(r->ve_ptr.Upload(ptr, c->stream));
      XRT_CHECK_HIP(r->ve_edge.Upload(adj, c->stream));
      XRT_CHECK_HIP(r->ee_gl.Alloc(std::max(ne, 1)));
    }
    A.rec[i] = std::move(r);
  }
  if (A.gx0.n != N) {
    XRT_CHECK_HIP(A.gx0.Alloc(N));
    XRT_CHECK_HIP(A.gv.Alloc(N));
    XRT_CHECK_HIP(A.gx.Alloc(N));
    XRT_CHECK_HIP(A.ee_cnt.Alloc(N));
    XRT_CHECK_HIP(hipMemsetAsync(A.ee_cnt.ptr, 0, sizeof(int32_t) * N, c->stream));
  }
  if (A.glam.n != static_cast<size_t>(std::max(A.e_total, 1))) {
    XRT_CHECK_HIP(A.glam.Alloc(std::max(A.e_total, 1)));
    XRT_CHECK_HIP(A.gcomp.Alloc(std::max(A.e_total, 1)));
  }
  if (A.gbcomp.n != static_cast<size_t>(std::max(A.b_total, 1)))
    XRT_CHECK_HIP(A.gbcomp.Alloc(std::max(A.b_total, 1)));
  if (A.rep) {
    const size_t marks = 2 * static_cast<size_t>(A.slots) * p.num_substeps * std::max(1, p.num_iterations);
    if (A.rep_pool.n != A.slots * A.rep_cap) XRT_CHECK_HIP(A.rep_pool.Alloc(A.slots * A.rep_cap));
    if (A.rep_count.n != static_cast<size_t>(A.slots)) XRT_CHECK_HIP(A.rep_count.Alloc(A.slots));
    if (A.rep_marks.n != marks) XRT_CHECK_HIP(A.rep_marks.Alloc(marks));
    if (A.rep_flags.n != 1) XRT_CHECK_HIP(A.rep_flags.Alloc(1));
    if (A.rep_cnt.n != N) {
      XRT_CHECK_HIP(A.rep_cnt.Alloc(N));
      XRT_CHECK_HIP(A.rep_add.Alloc(6 * N));
    }
    XRT_CHECK_HIP(hipMemsetAsync(A.rep_marks.ptr, 0, sizeof(unsigned long long) * marks, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(A.rep_flags.ptr, 0, sizeof(int32_t), c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(A.rep_cnt.ptr, 0, sizeof(int32_t) * N, c->stream));
    XRT_CHECK_HIP(hipMemsetAsync(A.rep_add.ptr, 0, sizeof(unsigned long long) * 6 * N, c->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(c->stream));
  }
  return XRT_OK;
}

int ZeroAdjoints(xrt_ctx* c) {
  AdjointData& A = *c->adj;
  XRT_CHECK_HIP(LaunchZeroD4(c->stream, A.gx0.ptr, A.n_total));
  XRT_CHECK_HIP(LaunchZeroD4(c->stream, A.gv.ptr, A.n_total));
  XRT_CHECK_HIP(LaunchZeroD1(c->stream, A.gcomp.ptr, static_cast<int>(A.gcomp.n)));
  XRT_CHECK_HIP(LaunchZeroD1(c->stream, A.gbcomp.ptr, static_cast<int>(A.gbcomp.n)));
  return XRT_OK;
}

int SaveState(xrt_ctx* c, float4* dst) {
  const int N = c->adj->n_total;
  XRT_CHECK_HIP(LaunchCopy4(c->stream, dst, c->x0.ptr, N));
  XRT_CHECK_HIP(LaunchCopy4(c->stream, dst + N, c->v.ptr, N));
  return XRT_OK;
}

int LoadState(xrt_ctx* c, const float4* src) {
  const int N = c->adj->n_total;
  XRT_CHECK_HIP(LaunchCopy4(c->stream, c->x0.ptr, src, N));
  XRT_CHECK_HIP(LaunchCopy4(c->stream, c->x.ptr, src, N));
  XRT_CHECK_HIP(LaunchCopy4(c->stream, c->v.ptr, src + N, N));
  return XRT_OK;
}

// contact records of pass p of the current frame -> the records of the current frame slot
int SnapshotRecords(xrt_ctx* c, int i, int frame_pass) {
  AdjointData& A = *c->adj;
  const HostInstance& h = c->inst[i];
  const ContactData& cd = *c->contact[i];
  AdjointData::Records& r = *A.rec[i];
  hipStream_t s = c->stream;
  XRT_REQUIRE(frame_pass >= 0 && frame_pass < A.n_passes, XRT_ERR_STATE,
              "collision pass index out of range");
  const size_t pass = static_cast<size_t>(A.slot) * A.n_passes + frame_pass;
  const size_t nv = h.nv, ne = static_cast<size_t>(std::max(h.ne(), 1));
  XRT_CHECK_HIP(hipMemcpyAsync(r.vf_face.ptr + pass * nv, cd.vf_face.ptr, 4 * nv, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(r.rt_face.ptr + pass * nv, cd.rt_face.ptr, 4 * nv, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(r.vf_q.ptr + pass * nv, cd.vf_q.ptr, 16 * nv, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(r.rt_q.ptr + pass * nv, cd.rt_q.ptr, 16 * nv, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(r.pred.ptr + pass * static_cast<size_t>(cd.obs_nv), cd.obs_bvh.pred.ptr,
                               16 * static_cast<size_t>(cd.obs_nv), hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(r.ee.ptr + pass * ne * sizeof(EeRec), cd.ee.ptr, ne * sizeof(EeRec),
                               hipMemcpyDeviceToDevice, s));
  return XRT_OK;
}

}  // namespace

// One substep of every instance as wide launches, the contact phases included; with a
// tape owner and sub >= 0 the operators' input states go to tape slot `sub`.
int WideSubstep(xrt_ctx* c, const StepConsts& k, const BendMode& bm, AdjointData* A, int sub) {
  const WideConfig& W = c->wide;
  hipStream_t s = c->stream;
  const int N = W.n_total, n_inst = static_cast<int>(c->inst.size());
  const WideArgs wa = MakeWideArgs(c);
  float4* T = (A && sub >= 0)
                  ? A->tape.ptr + (static_cast<size_t>(A->slot) * c->params.num_substeps + sub) *
                                      A->sub_stride
                  : nullptr;
  bool obs_contact = false;
  for (int i = 0; i < n_inst; ++i) obs_contact = obs_contact || ObsContact(c, i);
  int rc;
  if ((rc = MultiContactSubstepBegin(c)) != XRT_OK) return rc;
  if (T) XRT_CHECK_HIP(LaunchCopy4(s, T, c->x0.ptr, N));
  if (T && A->rep && sub == 0)  // a frame slot's response pool starts empty
    XRT_CHECK_HIP(hipMemsetAsync(A->rep_count.ptr + A->slot, 0, sizeof(unsigned long long), s));
  XRT_CHECK_HIP(LaunchPredict(s, c->x0.ptr, c->v.ptr, c->x.ptr, N, k));
  if ((rc = ParticleSubstepBegin(c, -1)) != XRT_OK) return rc;
  for (int it = 0; it < k.iterations; ++it) {
    float4* TI = T ? T + N + static_cast<size_t>(it) * A->it_stride : nullptr;
    if ((rc = ParticleIter(c, -1)) != XRT_OK) return rc;
    bool ran = false;
    RepTapeHook hook;
    if (TI && A->rep) {
      hook.x_out = TI + A->off_r;
      hook.pool = A->rep_pool.ptr + static_cast<size_t>(A->slot) * A->rep_cap;
      hook.count = A->rep_count.ptr + A->slot;
      hook.cap = A->rep_cap;
      hook.flags = A->rep_flags.ptr;
      hook.marks = A->rep_marks.ptr +
                   2 * ((static_cast<size_t>(A->slot) * c->params.num_substeps + sub) * k.iterations + it);
    }
    if ((rc = MultiContactIterPre(c, it, &ran, TI && A->rep ? &hook : nullptr)) != XRT_OK) return rc;
    if (TI && ran)
      for (int i = 0; i < n_inst; ++i)
        if (ObsContact(c, i) && (rc = SnapshotRecords(c, i, c->contact[i]->pass_cnt - 1)) != XRT_OK)
          return rc;
    if (FemOn(c)) {  // never taped: the adjoint refuses the FEM constraints
      if ((rc = FemIter(c, -1, it)) != XRT_OK) return rc;
    } else {
      for (int col = 0; col < static_cast<int>(W.s_max.size()); ++col)
        XRT_CHECK_HIP(LaunchWideStretch(s, wa, n_inst, col, W.s_max[col], k.inv_dt2, it,
                                        TI ? TI + A->off_s : nullptr, W.maps.ptr + W.s_map[col],
                                        W.s_blocks[col], c->tuning.wide_per_thread));
    }
    if (k.solve_bending && !FemOn(c))
      for (int col = 0; col < static_cast<int>(W.b_max.size()); ++col)
        XRT_CHECK_HIP(LaunchWideBend(s, wa, n_inst, col, W.b_max[col], k.inv_dt2, bm,
                                     TI ? TI + A->off_b : nullptr, W.maps.ptr + W.b_map[col],
                                     W.b_blocks[col], c->tuning.wide_per_thread));
    if (c->total_slots > 0) {
      if (TI) XRT_CHECK_HIP(LaunchCopy4(s, TI + A->off_l, c->x.ptr, N));
      XRT_CHECK_HIP(LaunchWideLra(s, c->desc.ptr, n_inst, c->max_slots, c->max_nv, c->x.ptr,
                                  c->slot_vert.ptr, c->slot_pos_all.ptr, c->lra_ptr.ptr,
                                  c->lra_slot.ptr, c->lra_rest.ptr, k.stretchiness));
    }
    if (k.enable_ground) {
      if (TI) XRT_CHECK_HIP(LaunchCopy4(s, TI + A->off_g, c->x.ptr, N));
      XRT_CHECK_HIP(LaunchGround(s, c->x.ptr, c->x0.ptr, N, k.ground_y, k.friction));
    }
    if (obs_contact) {
      if (TI) XRT_CHECK_HIP(LaunchCopy4(s, TI + A->off_c, c->x.ptr, N));
      if ((rc = MultiContactVfSolve(c)) != XRT_OK) return rc;
      if (TI) XRT_CHECK_HIP(LaunchCopy4(s, TI + A->off_e, c->x.ptr, N));
      if ((rc = MultiContactEeSolve(c)) != XRT_OK) return rc;
    }
    if ((rc = BindingIter(c, -1)) != XRT_OK) return rc;  // never taped: the adjoint refuses
  }
  XRT_CHECK_HIP(LaunchFinalize(s, c->x0.ptr, c->v.ptr, c->x.ptr, N, k));
  if ((rc = ParticleSubstepEnd(c, -1)) != XRT_OK) return rc;
  return MultiContactSubstepEnd(c);
}

namespace {

int WideFrame(xrt_ctx* c, bool tape) {
  const StepConsts k = MakeConsts(c->params);
  const BendMode bm = ModeOf(k);
  for (int sub = 0; sub < c->params.num_substeps; ++sub) {
    const int rc = WideSubstep(c, k, bm, c->adj.get(), tape ? sub : -1);
    if (rc != XRT_OK) return rc;
  }
  return XRT_OK;
}

// Reverse of the substep in tape slot `sub` (its position in the frame): the adjoints of
// its end state (gx0, gv) become those of its start state.
int ReverseSubstep(xrt_ctx* c, const StepConsts& k, const BendMode& bm, int sub) {
  AdjointData& A = *c->adj;
  hipStream_t s = c->stream;
  const int N = A.n_total, n_inst = static_cast<int>(c->inst.size());
  const WideArgs wa = MakeWideArgs(c);
  const float4* T =
      A.tape.ptr + (static_cast<size_t>(A.slot) * c->params.num_substeps + sub) * A.sub_stride;
  // the forward's finalize: v = (x - x0) * (1 / dt) * damp_factor (FP32 factors)
  const double fin = static_cast<double>(1.0f / k.dt) * static_cast<double>(k.damp_factor);
  XRT_CHECK_HIP(LaunchFinalizeAdj(s, A.gx0.ptr, A.gv.ptr, A.gx.ptr, N, fin));
  XRT_CHECK_HIP(LaunchZeroD1(s, A.glam.ptr, static_cast<int>(A.glam.n)));
  for (int it = k.iterations - 1; it >= 0; --it) {
    const float4* TI = T + N + static_cast<size_t>(it) * A.it_stride;
    if (A.contacts) {
      const int pass = A.slot * A.n_passes + (sub * k.iterations + it) / A.frequency;
      for (int i = 0; i < n_inst; ++i) {
        if (!ObsContact(c, i)) continue;
        const HostInstance& h = c->inst[i];
        const ContactData& cd = *c->contact[i];
        const AdjointData::Records& r = *A.rec[i];
        const int vb = h.desc.vbeg, nv = h.nv;
        const size_t pv = static_cast<size_t>(pass) * nv;
        const EeRec* ee = reinterpret_cast<const EeRec*>(r.ee.ptr) +
                          static_cast<size_t>(pass) * std::max(h.ne(), 1);
        XRT_CHECK_HIP(LaunchEeAdj(s, TI + A.off_e + vb, cd.c_edges.ptr, h.ne(), nv, ee,
                                  c->params.relaxation_factor, A.gx.ptr + vb, A.ee_cnt.ptr + vb,
                                  r.ee_gl.ptr, r.ve_ptr.ptr, r.ve_edge.ptr));
        XRT_CHECK_HIP(LaunchVfAdj(s, TI + A.off_c + vb, nv, cd.o_tris.ptr,
                                  r.pred.ptr + static_cast<size_t>(pass) * cd.obs_nv,
                                  r.rt_face.ptr + pv, r.rt_q.ptr + pv, r.vf_face.ptr + pv,
                                  r.vf_q.ptr + pv, A.gx.ptr + vb));
      }
    }
    if (k.enable_ground)
      XRT_CHECK_HIP(LaunchGroundAdj(s, TI + A.off_g, T, A.gx.ptr, A.gx0.ptr, N, k.ground_y, k.friction));
    if (A.lra)
      XRT_CHECK_HIP(LaunchWideLraAdj(s, c->desc.ptr, n_inst, c->max_nv, c->total_slots,
                                     TI + A.off_l, c->slot_vert.ptr, c->lra_ptr.ptr,
                                     c->lra_slot.ptr, c->lra_rest.ptr, k.stretchiness,
                                     A.lra_xs.ptr, A.lra_gslot.ptr, A.slot_inst.ptr,
                                     A.slot_ptr.ptr, A.slot_cons.ptr, A.gx.ptr));
    if (k.solve_bending)
      for (int col = static_cast<int>(A.b_max.size()) - 1; col >= 0; --col)
        XRT_CHECK_HIP(LaunchWideBendAdj(s, wa, n_inst, col, A.b_max[col], k.inv_dt2, bm,
                                        TI + A.off_b, A.gx.ptr, A.gbcomp.ptr));
    for (int col = static_cast<int>(A.s_max.size()) - 1; col >= 0; --col)
      XRT_CHECK_HIP(LaunchWideStretchAdj(s, wa, n_inst, col, A.s_max[col], k.inv_dt2, it,
                                         TI + A.off_s, A.gx.ptr, A.glam.ptr, A.gcomp.ptr));
    if (A.rep) {  // the iteration's repulsion pass came first
      const size_t mi = (static_cast<size_t>(A.slot) * c->params.num_substeps + sub) * k.iterations + it;
      XRT_CHECK_HIP(RepulsionAdj(s, c->ic.ptr, N, A.rep_pool.ptr + static_cast<size_t>(A.slot) * A.rep_cap,
                                 A.rep_marks.ptr + 2 * mi, A.rep_cap, TI + A.off_r, T,
                                 c->params.pbd_thickness, c->params.repulsion_relaxation,
                                 A.rep_cnt.ptr, A.gx.ptr, A.gx0.ptr, A.rep_add.ptr,
                                 A.rep_add.ptr + 3 * static_cast<size_t>(N)));
    }
  }
  XRT_CHECK_HIP(LaunchPredictAdj(s, A.gx0.ptr, A.gv.ptr, A.gx.ptr, N, k.dt));
  return XRT_OK;
}

// contact bookkeeping back to a frame start (after a substep that is not kept)
void RewindContacts(xrt_ctx* c) {
  for (auto& cd : c->contact)
    if (cd) {
      cd->frame_substep = 0;
      cd->pass_cnt = 0;
    }
}

int CheckAdj(xrt_ctx* ctx, int32_t inst) {
  XRT_REQUIRE(ctx != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(ctx->adj && ctx->adj->recorded, XRT_ERR_STATE, "xrt_adjoint_begin has not been called");
  XRT_REQUIRE(inst >= -1 && inst < static_cast<int32_t>(ctx->inst.size()), XRT_ERR_INVALID_ARG,
              "instance index out of range");
  return XRT_OK;
}

}  // namespace
}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_adjoint_begin(xrt_ctx* ctx, int32_t nframes, int32_t checkpoint_every) {
  XRT_REQUIRE(ctx != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(nframes > 0 && checkpoint_every > 0, XRT_ERR_INVALID_ARG,
              "nframes and checkpoint_every must be > 0");
  int rc = AdjSetup(ctx, std::min(checkpoint_every, nframes));
  if (rc != XRT_OK) return rc;
  AdjointData& A = *ctx->adj;
  A.recorded = false;
  A.has_grad = false;
  A.n_frames = nframes;
  A.ckpt_every = std::min(checkpoint_every, nframes);
  const size_t N = static_cast<size_t>(A.n_total);
  const int n_ckpt = (nframes + A.ckpt_every - 1) / A.ckpt_every;
  XRT_CHECK_HIP(A.ckpt.Alloc(static_cast<size_t>(n_ckpt) * 2 * N));
  XRT_CHECK_HIP(A.fstart.Alloc(static_cast<size_t>(A.ckpt_every) * 2 * N));
  XRT_CHECK_HIP(A.final_state.Alloc(2 * N));
  for (int f = 0; f < nframes; ++f) {
    if (f % A.ckpt_every == 0 &&
        (rc = SaveState(ctx, A.ckpt.ptr + static_cast<size_t>(f / A.ckpt_every) * 2 * N)) != XRT_OK)
      return rc;
    if ((rc = WideFrame(ctx, false)) != XRT_OK) return rc;
  }
  if ((rc = SaveState(ctx, A.final_state.ptr)) != XRT_OK) return rc;
  if ((rc = ZeroAdjoints(ctx)) != XRT_OK) return rc;
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  A.recorded = true;
  return XRT_OK;
}

int xrt_adjoint_seed(xrt_ctx* ctx, int32_t inst, const float* grad_x, const float* grad_v) {
  int rc = CheckAdj(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(inst >= 0, XRT_ERR_INVALID_ARG, "instance index out of range");
  const HostInstance& h = ctx->inst[inst];
  AdjointData& A = *ctx->adj;
  if (A.has_grad) {  // a new seed after a reverse pass starts from zero seeds
    if ((rc = ZeroAdjoints(ctx)) != XRT_OK) return rc;
    A.has_grad = false;
  }
  std::vector<double4> buf(h.nv);
  for (int which = 0; which < 2; ++which) {
    const float* g = which == 0 ? grad_x : grad_v;
    for (int i = 0; i < h.nv; ++i) {
      buf[i] = g ? make_double4(g[3 * i], g[3 * i + 1], g[3 * i + 2], 0.0) : make_double4(0, 0, 0, 0);
      XRT_REQUIRE(std::isfinite(buf[i].x) && std::isfinite(buf[i].y) && std::isfinite(buf[i].z),
                  XRT_ERR_NOT_FINITE, "non-finite adjoint seed");
    }
    double4* dst = (which == 0 ? A.gx0.ptr : A.gv.ptr) + h.desc.vbeg;
    XRT_CHECK_HIP(hipMemcpyAsync(dst, buf.data(), sizeof(double4) * h.nv, hipMemcpyHostToDevice,
                                 ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  return XRT_OK;
}

// after a taped forward: the repulsion responses fitted their pool and every instance's
// repulsion ran from its item lists (the per-pair fallback has no replay)
int CheckRepTape(xrt_ctx* ctx) {
  AdjointData& A = *ctx->adj;
  if (!A.rep) return XRT_OK;
  int32_t f = 0;
  XRT_CHECK_HIP(hipMemcpyAsync(&f, A.rep_flags.ptr, sizeof(f), hipMemcpyDeviceToHost, ctx->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  XRT_REQUIRE(!(f & 1), XRT_ERR_CAPACITY,
              "the repulsion responses of a frame exceeded the adjoint's record pool ("
              "8 per node and pass, at most 2^27)");
  XRT_REQUIRE(!(f & 2), XRT_ERR_UNSUPPORTED,
              "an instance's repulsion items exceeded its item lists (repulsion_vf_items_per_face / "
              "repulsion_ee_items_per_face, or the pair capacity): the per-pair evaluation has no "
              "reverse; raise the capacities");
  return XRT_OK;
}

int xrt_adjoint_backward(xrt_ctx* ctx) {
  int rc = CheckAdj(ctx, -1);
  if (rc != XRT_OK) return rc;
  AdjointData& A = *ctx->adj;
  XRT_REQUIRE(!A.has_grad, XRT_ERR_STATE,
              "the seeds were consumed by the previous reverse pass: seed again first");
  XRT_REQUIRE(!ctx->params.enable_particle_collision &&
                  (ctx->params.enable_self_contact != 0) == A.rep && !FemOn(ctx),
              XRT_ERR_UNSUPPORTED,
              "particle self-collision / self contact / FEM constraints changed after "
              "xrt_adjoint_begin");
  const StepConsts k = MakeConsts(ctx->params);
  const BendMode bm = ModeOf(k);
  const size_t N = static_cast<size_t>(A.n_total);
  const int K = A.ckpt_every, n_ckpt = (A.n_frames + K - 1) / K;
  XRT_CHECK_HIP(LaunchZeroD1(ctx->stream, A.gcomp.ptr, static_cast<int>(A.gcomp.n)));
  XRT_CHECK_HIP(LaunchZeroD1(ctx->stream, A.gbcomp.ptr, static_cast<int>(A.gbcomp.n)));
  if (A.rep) XRT_CHECK_HIP(hipMemsetAsync(A.rep_flags.ptr, 0, sizeof(int32_t), ctx->stream));
  for (int seg = n_ckpt - 1; seg >= 0; --seg) {
    const int f0 = seg * K, f1 = std::min(f0 + K, A.n_frames);
    if ((rc = LoadState(ctx, A.ckpt.ptr + static_cast<size_t>(seg) * 2 * N)) != XRT_OK) return rc;
    if (A.slots >= f1 - f0) {
      // one taped forward over the segment (frame f in slot f - f0), then its reverse
      RewindContacts(ctx);
      for (int f = f0; f < f1; ++f) {
        A.slot = f - f0;
        if ((rc = WideFrame(ctx, true)) != XRT_OK) return rc;
      }
      if ((rc = CheckRepTape(ctx)) != XRT_OK) return rc;
      for (int f = f1 - 1; f >= f0; --f) {
        A.slot = f - f0;
        for (int sub = ctx->params.num_substeps - 1; sub >= 0; --sub)
          if ((rc = ReverseSubstep(ctx, k, bm, sub)) != XRT_OK) return rc;
      }
      A.slot = 0;
      continue;
    }
    // one frame slot: frame starts from an untaped pass, then each frame taped and reversed
    for (int f = f0; f < f1; ++f) {
      if ((rc = SaveState(ctx, A.fstart.ptr + static_cast<size_t>(f - f0) * 2 * N)) != XRT_OK) return rc;
      if (f + 1 < f1 && (rc = WideFrame(ctx, false)) != XRT_OK) return rc;
    }
    RewindContacts(ctx);
    for (int f = f1 - 1; f >= f0; --f) {
      if ((rc = LoadState(ctx, A.fstart.ptr + static_cast<size_t>(f - f0) * 2 * N)) != XRT_OK) return rc;
      if ((rc = WideFrame(ctx, true)) != XRT_OK) return rc;
      if ((rc = CheckRepTape(ctx)) != XRT_OK) return rc;
      for (int sub = ctx->params.num_substeps - 1; sub >= 0; --sub)
        if ((rc = ReverseSubstep(ctx, k, bm, sub)) != XRT_OK) return rc;
    }
  }
  if ((rc = LoadState(ctx, A.final_state.ptr)) != XRT_OK) return rc;
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  A.has_grad = true;
  return XRT_OK;
}

int xrt_adjoint_get_grad_compliance(xrt_ctx* ctx, int32_t inst, float* grad_stretch,
                                    float* grad_bend) {
  int rc = CheckAdj(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(inst >= 0, XRT_ERR_INVALID_ARG, "instance index out of range");
  const HostInstance& h = ctx->inst[inst];
  AdjointData& A = *ctx->adj;
  for (int which = 0; which < 2; ++which) {
    float* out = which == 0 ? grad_stretch : grad_bend;
    const int n = which == 0 ? h.ne() : h.nb();
    if (!out || n == 0) continue;
    std::vector<double> g(n);
    const double* src = (which == 0 ? A.gcomp.ptr + h.sbeg : A.gbcomp.ptr + h.bbeg);
    XRT_CHECK_HIP(hipMemcpyAsync(g.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    const std::vector<int32_t>& order = which == 0 ? h.s_order : h.b_order;
    for (int p = 0; p < n; ++p) out[order[p]] = static_cast<float>(g[p]);
  }
  return XRT_OK;
}

int xrt_adjoint_get_grad_state(xrt_ctx* ctx, int32_t inst, float* grad_x0, float* grad_v0) {
  int rc = CheckAdj(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(inst >= 0, XRT_ERR_INVALID_ARG, "instance index out of range");
  const HostInstance& h = ctx->inst[inst];
  AdjointData& A = *ctx->adj;
  std::vector<double4> buf(h.nv);
  for (int which = 0; which < 2; ++which) {
    float* out = which == 0 ? grad_x0 : grad_v0;
    if (!out) continue;
    XRT_CHECK_HIP(hipMemcpyAsync(buf.data(), (which == 0 ? A.gx0.ptr : A.gv.ptr) + h.desc.vbeg,
                                 sizeof(double4) * h.nv, hipMemcpyDeviceToHost, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < h.nv; ++i) {
      out[3 * i] = static_cast<float>(buf[i].x);
      out[3 * i + 1] = static_cast<float>(buf[i].y);
      out[3 * i + 2] = static_cast<float>(buf[i].z);
    }
  }
  return XRT_OK;
}

int xrt_adjoint_substep(xrt_ctx* ctx) {
  XRT_REQUIRE(ctx != nullptr, XRT_ERR_INVALID_ARG, "null context");
  const bool had = ctx->adj && ctx->adj->gx0.n > 0;
  int rc = AdjSetup(ctx);
  if (rc != XRT_OK) return rc;
  AdjointData& A = *ctx->adj;
  if (!had && (rc = ZeroAdjoints(ctx)) != XRT_OK) return rc;
  const StepConsts k = MakeConsts(ctx->params);
  const BendMode bm = ModeOf(k);
  if (A.final_state.n != 2 * static_cast<size_t>(A.n_total))
    XRT_CHECK_HIP(A.final_state.Alloc(2 * static_cast<size_t>(A.n_total)));
  if ((rc = SaveState(ctx, A.final_state.ptr)) != XRT_OK) return rc;
  if ((rc = WideSubstep(ctx, k, bm, &A, 0)) != XRT_OK) return rc;
  RewindContacts(ctx);
  if ((rc = LoadState(ctx, A.final_state.ptr)) != XRT_OK) return rc;
  if ((rc = ReverseSubstep(ctx, k, bm, 0)) != XRT_OK) return rc;
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  A.recorded = true;  // gradients readable
  A.has_grad = true;
  return XRT_OK;
}

int xrt_time_adjoint(xrt_ctx* ctx, int32_t n, float ms[3]) {
  XRT_REQUIRE(ctx && ms && n > 0, XRT_ERR_INVALID_ARG, "bad argument");
  int rc = AdjSetup(ctx);
  if (rc != XRT_OK) return rc;
  AdjointData& A = *ctx->adj;
  const StepConsts k = MakeConsts(ctx->params);
  const BendMode bm = ModeOf(k);
  const size_t N = static_cast<size_t>(A.n_total);
  DevBuf<float4> keep;
  XRT_CHECK_HIP(keep.Alloc(2 * N));
  if ((rc = SaveState(ctx, keep.ptr)) != XRT_OK) return rc;
  if ((rc = ZeroAdjoints(ctx)) != XRT_OK) return rc;
  const int S = ctx->params.num_substeps;
  for (int phase = 0; phase < 3; ++phase) {
    if ((rc = LoadState(ctx, keep.ptr)) != XRT_OK) return rc;
    RewindContacts(ctx);
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    XRT_CHECK_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int j = 0; j < n; ++j) {
      if (phase == 0) rc = WideSubstep(ctx, k, bm, nullptr, -1);
      else if (phase == 1) rc = WideSubstep(ctx, k, bm, &A, j % S);
      else rc = ReverseSubstep(ctx, k, bm, (n - 1 - j) % S);
      if (rc != XRT_OK) return rc;
    }
    XRT_CHECK_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    XRT_CHECK_HIP(hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    XRT_CHECK_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    ms[phase] = t / static_cast<float>(n);
  }
  if ((rc = LoadState(ctx, keep.ptr)) != XRT_OK) return rc;
  RewindContacts(ctx);
  if ((rc = ZeroAdjoints(ctx)) != XRT_OK) return rc;
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  A.recorded = false;
  A.has_grad = false;
  return XRT_OK;
}

}  // extern "C"
