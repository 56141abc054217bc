// Solver context: host mirror of every instance plus the device buffers (one HIP stream).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/xrtailor_amd.h"
#include "binding.h"
#include "blockmap.h"
#include "bvh.h"
#include "common.h"
#include "contact.h"
#include "device.h"
#include "fem.h"
#include "particle.h"
#include "topology.h"

namespace xrt {

struct HostInstance {
  int nv = 0, nt = 0;
  std::vector<float> rest_pos;  // 3*nv
  std::vector<int32_t> tris;    // 3*nt
  std::vector<float> inv_mass;  // nv
  std::vector<int32_t> edges;   // 2*ne, original order
  std::vector<float> rest_len, s_comp;
  std::vector<int32_t> bends;  // 4*nb
  std::vector<float> theta, b_comp;
  std::vector<int32_t> slots;  // LRA slot vertices
  std::vector<std::vector<double>> slot_dist;  // geodesic distance per slot
  std::vector<float> x0_init, v_init;          // optional initial state (3*nv)
  // built by xrt_start
  std::vector<int32_t> s_color, b_color;
  int s_ncol = 0, b_ncol = 0;
  std::vector<int32_t> s_order, b_order;    // sorted position -> original index
  std::vector<int32_t> s_coloff, b_coloff;  // local color offsets (ncol + 1)
  std::vector<int32_t> lra_ptr, lra_slot;
  std::vector<float> lra_rest;
  InstDesc desc{};
  int sbeg = 0, bbeg = 0, kbeg = 0;
  int ne() const { return static_cast<int>(edges.size() / 2); }
  int nb() const { return static_cast<int>(bends.size() / 4); }
  // obstacle (optional): rest/initial positions, faces, edges
  std::vector<float> obs_pos;
  std::vector<int32_t> obs_tris, obs_edges;
  int obs_nv() const { return static_cast<int>(obs_pos.size() / 3); }
  int obs_nt() const { return static_cast<int>(obs_tris.size() / 3); }
  ContactTopo cloth_topo, obs_topo;  // built by xrt_start
  // cloth-to-body bindings (xrt_add_bindings) and the obstacle's skinning weights
  // (xrt_set_obstacle_skinning)
  std::vector<int32_t> bind_vert;
  std::vector<float> bind_stiff, bind_dist;
  std::vector<int32_t> skin_ptr, skin_joint;
  std::vector<float> skin_weight;
  int n_joints = 0;
};

}  // namespace xrt

// Device-side contact state of one instance (plain path).
namespace xrt {
struct ContactData {
  // cloth topology
  DevBuf<int4> c_tris, c_fe;
  DevBuf<int2> c_edges;
  DevBuf<int32_t> c_ef_ptr, c_ef;
  DevBuf<float4> c_face_n;
  DevBuf<uint32_t> c_rtri;
  DevBuf<int4> c_trir;
  // obstacle topology and state
  int obs_nv = 0, obs_nf = 0, obs_ne = 0;
  DevBuf<int4> o_tris, o_fe;
  DevBuf<int2> o_edges;
  DevBuf<int32_t> o_ef_ptr, o_ef;
  DevBuf<float4> o_face_n;
  DevBuf<uint32_t> o_rtri;
  DevBuf<int4> o_trir;
  DevBuf<float4> o_x0, o_x, o_x1, o_n1;
  DevBuf<int32_t> o_vptr, o_corner;  // vertex -> corner CSR for normals
  // trees
  Lbvh obs_bvh, cloth_bvh;
  // contact records and Jacobi accumulators
  DevBuf<int32_t> vf_face, rt_face;
  DevBuf<float4> vf_q, rt_q;
  DevBuf<char> ee;  // EeRec[E]
  DevBuf<uint8_t> ee_none;  // per edge: 1 when its EE record is empty (a.x > 1e6)
  DevBuf<long long> dq;
  DevBuf<int32_t> cnt;
  // face pairs of the repulsion (cloth face, other face)
  int pair_cap = 0;
  DevBuf<int2> pairs;
  DevBuf<int32_t> scal;  // [0] zero, [1] obstacle-pair count, [2] total pair count,
                         // [3..8] repulsion item counts (batched path, InstContactDev)
  // repulsion test items of the batched path (capacities 0: lists off)
  int vf_cap = 0, ee_cap = 0;
  DevBuf<int2> rep_vf, rep_ee;
  DevBuf<uint32_t> ivf, iee;  // per cloth leaf item offsets (block-local prefix)
  DevBuf<uint2> nitem;        // per cloth leaf item counts (VF, EE)
  DevBuf<uint32_t> stage;  // kPairStage hits per cloth leaf of the batched pair refresh
  DevBuf<uint32_t> bsum;  // per-block sums of the batched pair refresh
  int frame_substep = 0, pass_cnt = 0;
  bool obstacle_moving = false;  // xrt_update_obstacle since the last frame end
};

// Particle self-collision state of every instance (particle.h has the HBM layout).
struct ParticleData {
  bool sized = false;      // buffers allocated for max_nb
  int max_nb = 0, passes = 0;
  std::vector<ParticleInstDev> inst;
  std::vector<int> substep;  // substep of the current frame, per instance
  DevBuf<ParticleInstDev> dev;
  DevBuf<uint32_t> keys, keys_tmp, cell_start, cell_end, nb, hist;
  DevBuf<int32_t> vals, vals_tmp;
  DevBuf<float4> orig, delta;
  DevBuf<int32_t> maps;  // block maps: every instance (vertex, tile), then each instance's
  MapRef all_v, all_t;
  std::vector<MapRef> one_v, one_t;
  DevBuf<unsigned long long> count;  // list-entry counter of xrt_time_particles
  // xrt_time_particles: event pairs around the collision kernels and the hash steps
  std::vector<hipEvent_t> collide_events, hash_events;
  size_t collide_next = 0, hash_next = 0;
};

// Bindings, nearest-face records and obstacle skinning of every instance (binding.h).
// Pinned two-slot staging of the joint transforms (xrt_skin_obstacle[s]): the upload is
// queued on the stream without waiting for it; a slot is written again only after the copy
// issued from it two calls earlier has completed.
struct SkinStage {
  float* host = nullptr;  // 2 slots of `per` floats
  size_t per = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool fresh[2] = {true, true};
  int slot = 0;
  SkinStage() = default;
  SkinStage(const SkinStage&) = delete;
  SkinStage& operator=(const SkinStage&) = delete;
  void Release();
  ~SkinStage() { Release(); }
};

struct BindingData {
  int total = 0;                   // bindings over every instance
  std::vector<BindInstDev> inst;   // per instance (nb 0 when it has none)
  DevBuf<BindInstDev> dev;
  DevBuf<BindItem> items;
  DevBuf<int4> near_idx;           // nearest-face record of every cloth vertex
  DevBuf<float4> near_uvw;
  DevBuf<int32_t> flag;
  DevBuf<int32_t> maps;            // binding blocks: every instance, then each alone
  MapRef all;
  std::vector<MapRef> one;
  // skinning
  int total_joints = 0;
  std::vector<int> joint_off;      // per instance: offset of its transforms (-1: none)
  std::vector<SkinInstDev> skin;
  DevBuf<SkinInstDev> skin_dev;
  DevBuf<int32_t> sk_ptr, sk_joint;
  DevBuf<float> sk_weight, transforms;
  DevBuf<float4> sk_rest;
  DevBuf<int32_t> skin_maps;       // obstacle-vertex blocks: every instance, each alone
  MapRef skin_all;
  std::vector<MapRef> skin_one;
  SkinStage stage;
};

// Quality-mode FEM constraints of every instance (fem.h has the HBM layout).  Built at
// xrt_start, or at xrt_set_params when the constraint model turns to FEM later.
struct FemData {
  bool ready = false;
  uint64_t seed = 0;  // xrt_start's coloring seed
  struct Inst {
    std::vector<int32_t> color, order, coloff;  // face colors, sorted -> face, local offsets
    int ncol = 0, fbeg = 0;
  };
  std::vector<Inst> inst;
  DevBuf<StrainC> sc;            // faces in color order per instance
  DevBuf<int32_t> f_coloff;      // global color offsets (ncol + 1 per instance)
  DevBuf<int2> fcol;             // per instance: first offset, colors
  DevBuf<float4> iso_k;          // isometric bending, in the bend elements' order
  DevBuf<float> iso_c, iso_lam;
  std::vector<size_t> f_map;     // wide strain launches: per-color block maps in `maps`
  std::vector<int> f_blocks;
  DevBuf<int32_t> maps;
};

// Per-frame cache of xrt_record_frames (record_api.hip): two device slots packed after
// each frame, copied on a second stream into two pinned host slots.
struct RecordData {
  hipStream_t copy_stream = nullptr;
  hipEvent_t packed[2] = {nullptr, nullptr}, copied[2] = {nullptr, nullptr};
  bool fresh[2] = {true, true};  // slot never copied yet (nothing to wait for)
  size_t slot_floats = 0;
  DevBuf<float> dev;
  float* host = nullptr;  // pinned, 2 slots
  DevBuf<int4> tris;      // every instance's faces, global vertex ids (normals)
  DevBuf<int32_t> vptr, corner;
  DevBuf<float4> normals;
  RecordData() = default;
  RecordData(const RecordData&) = delete;
  RecordData& operator=(const RecordData&) = delete;
  void Release();
  ~RecordData() { Release(); }
};

// Sizes of the wide launches (one color of every instance per launch).
struct WideConfig {
  int n_total = 0, e_total = 0, b_total = 0;  // vertices / stretch / bend over instances
  std::vector<int> s_max, b_max;              // largest color c over instances
  // per color: offset of its block map in `maps` (n + 1 prefix sums of the instances'
  // 256-thread block counts, then the instance of every block) and its block total
  std::vector<size_t> s_map, b_map;
  std::vector<int> s_blocks, b_blocks;
  DevBuf<int32_t> maps;
};

// Adjoint (differentiable rollout) state: checkpoints, one frame of tape, adjoints.
struct AdjointData {
  int n_frames = 0, ckpt_every = 1, n_passes = 0, frequency = 1;
  bool recorded = false, has_grad = false, contacts = false;
  int n_total = 0, e_total = 0, b_total = 0;  // copies of the context's WideConfig
  std::vector<int> s_max, b_max;
  // tape of `slots` frames: frame slot f, substep s at (f * num_substeps + s) * sub_stride;
  // iteration it at n_total + it * it_stride within the substep.  `slot` is the frame slot
  // the taped forward writes and the reverse reads.
  size_t it_stride = 0, sub_stride = 0, off_s = 0, off_b = 0, off_l = 0, off_g = 0, off_c = 0,
         off_e = 0;
  int slots = 1, slot = 0;
  DevBuf<float4> tape;
  // contact records of each collision pass of each frame slot (slot * n_passes + pass), per
  // instance
  struct Records {
    DevBuf<int32_t> vf_face, rt_face;
    DevBuf<float4> vf_q, rt_q, pred;
    DevBuf<char> ee;
    // EE reverse: each vertex's edges (CSR, ascending) and a per-edge multiplier adjoint
    DevBuf<int32_t> ve_ptr, ve_edge;
    DevBuf<double> ee_gl;
  };
  std::vector<std::unique_ptr<Records>> rec;
  DevBuf<float4> ckpt;    // (x0, v) every ckpt_every frames: [n_ckpt][2][n_total]
  DevBuf<float4> fstart;  // (x0, v) at each frame start of a segment: [ckpt_every][2][n_total]
  DevBuf<float4> final_state;  // (x0, v) after the recorded frames
  DevBuf<double4> gx0, gv, gx;  // FP64 adjoints
  DevBuf<double> glam, gcomp, gbcomp;
  DevBuf<int32_t> ee_cnt;
  // PBD repulsion (self contact): per frame slot a pool of the responses of every repulsion
  // pass (instance, is_ee, item), per (slot, substep, iteration) their range (begin, end);
  // the state before each pass is taped at off_r
  bool rep = false;
  size_t off_r = 0;
  unsigned long long rep_cap = 0;        // pool entries per frame slot
  DevBuf<int4> rep_pool;                 // slots x rep_cap
  DevBuf<unsigned long long> rep_count;  // per slot
  DevBuf<unsigned long long> rep_marks;  // 2 per (slot, substep, iteration)
  DevBuf<int32_t> rep_flags, rep_cnt;    // 1; n_total
  DevBuf<unsigned long long> rep_add;    // 6 n_total (32.32 fixed point)
  // long-range attachments: per constraint, the target's position before it (replay) and
  // its slot's adjoint; per global slot, its instance and its constraints (in order)
  bool lra = false;
  DevBuf<float4> lra_xs;
  DevBuf<double4> lra_gslot;
  DevBuf<int32_t> slot_inst, slot_ptr, slot_cons;
};
}  // namespace xrt

struct xrt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  xrt_params params{};
  std::vector<xrt::HostInstance> inst;
  bool started = false;
  // device properties
  int cu_count = 0;
  size_t max_lds = 0;
  int wave_size = 64;
  int64_t l2_bytes = 0;
  // device buffers
  xrt::DevBuf<float4> x0, v, x;
  xrt::DevBuf<xrt::StretchC> sc;
  xrt::DevBuf<float> slam;
  xrt::DevBuf<int4> bidx;
  xrt::DevBuf<float2> bpar;
  xrt::DevBuf<int32_t> s_coloff, b_coloff;
  xrt::DevBuf<int32_t> slot_vert, lra_ptr, lra_slot;
  xrt::DevBuf<float4> slot_pos_all;  // slot snapshot of every instance (wide LRA)
  int total_slots = 0, max_nv = 0;
  xrt::DevBuf<float> lra_rest;
  xrt::DevBuf<float4> slot_pos;  // plain-path LRA snapshot
  xrt::DevBuf<xrt::InstDesc> desc;
  xrt::DevBuf<int32_t> order;
  // batched launch configuration
  // launch shapes picked by xrt_autotune (defaults until it runs); none changes a result
  xrt_tuning tuning{1024, 0, 1, 0, 0.0f, 0.0f, 0};
  size_t batched_lds = 0;
  int lds_verts = 0;
  int max_slots = 0;
  // per-instance contact state (allocated by xrt_start when contacts can occur)
  std::vector<std::unique_ptr<xrt::ContactData>> contact;
  // device views of the contact state and the block -> instance maps of the
  // multi-instance contact kernels
  xrt::DevBuf<xrt::InstContactDev> ic;
  xrt::DevBuf<int32_t> ic_off;
  // pair-overflow word (InstContactDev::overflow), read and cleared by CheckPairOverflow
  xrt::DevBuf<int32_t> overflow;
  xrt::TreeSet cloth_trees, obs_trees;  // frame-start builds of the batched path
  xrt::MultiMaps maps;
  // wide (all instances per launch) launch sizes
  xrt::WideConfig wide;
  // swift-mode particle self-collision
  xrt::ParticleData particles;
  // obstacle skinning and cloth-to-body bindings
  xrt::BindingData binding;
  // per-frame cache (xrt_record_frames)
  xrt::RecordData record;
  // quality-mode FEM constraints
  xrt::FemData fem;
  // differentiable rollout (allocated by xrt_adjoint_begin)
  std::unique_ptr<xrt::AdjointData> adj;
  // graph of one batched frame, and the frame schedule it was captured at (each contact
  // instance's substep of the frame and collision passes, each particle instance's substep)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<int> graph_key;
  // timing events
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // xrt_time_repulsion: event pairs recorded around the EE repulsion launches (next free pair)
  std::vector<hipEvent_t> rep_events;
  size_t rep_next = 0;
};

namespace xrt {
// api.hip
StepConsts MakeConsts(const xrt_params& p);
BendMode ModeOf(const StepConsts& k);
// sort constraints by (color, original index): order = sorted -> original, local offsets
void ColorSort(const std::vector<int32_t>& color, int ncol, std::vector<int32_t>* order,
               std::vector<int32_t>* coloff);
// contact_api.hip
bool ContactNeeded(const xrt_ctx* ctx, const HostInstance& h);
int ContactStart(xrt_ctx* ctx);
// pair / item lists re-allocated at the current capacities (after xrt_set_params changed
// them on a started context)
int ContactResizeLists(xrt_ctx* ctx);
// synchronizes the stream, then XRT_ERR_CAPACITY (and clears the word) if an instance's
// face pairs exceeded the pair capacity since the last check
int CheckPairOverflow(xrt_ctx* ctx);
// ApplyDeltas of one instance's contact Jacobi sums (no-op without contact state)
int ApplyInstanceDeltas(xrt_ctx* ctx, int32_t inst);
int SubstepPlainContact(xrt_ctx* ctx, int inst, const StepConsts& k, const BendMode& bm);
// contact phases of one instance's substep (shared by the plain and the wide drivers)
int ContactSubstepBegin(xrt_ctx* ctx, int inst);
int ContactIterPre(xrt_ctx* ctx, int inst, int it, bool* pass_ran);
int ContactVfSolve(xrt_ctx* ctx, int inst);
int ContactEeSolve(xrt_ctx* ctx, int inst);
int ContactSubstepEnd(xrt_ctx* ctx, int inst);
// after xrt_set_state on a started context: obstacle contact records cleared, cloth tree
// rebuilt from the new positions
int ContactStateReset(xrt_ctx* ctx, int inst);
// the same phases for every instance at once (multi-instance kernels)
bool ContactsActive(const xrt_ctx* ctx);
int MultiContactSetup(xrt_ctx* ctx);
int MultiContactSubstepBegin(xrt_ctx* ctx);
// with `rep_tape` (the adjoint's taped forward): the state before the repulsion pass is
// copied to rep_tape->x_out and the pass's responses recorded (MultiRepRecord)
struct RepTapeHook {
  float4* x_out = nullptr;
  int4* pool = nullptr;
  unsigned long long* count = nullptr;
  unsigned long long cap = 0;
  int32_t* flags = nullptr;
  unsigned long long* marks = nullptr;  // (begin, end)
};
int MultiContactIterPre(xrt_ctx* ctx, int it, bool* pass_ran, const RepTapeHook* rep_tape = nullptr);
int MultiContactVfSolve(xrt_ctx* ctx);
int MultiContactEeSolve(xrt_ctx* ctx);
int MultiContactSubstepEnd(xrt_ctx* ctx);
// particle_api.hip: swift-mode particle self-collision.  ParticleStart snapshots the
// initial positions (xrt_start); ParticleConfigure sizes the buffers for the current
// parameters.  The phases take an instance, or -1 for every instance at once.
bool ParticlesOn(const xrt_ctx* ctx);
int ParticleStart(xrt_ctx* ctx);
int ParticleConfigure(xrt_ctx* ctx);
int ParticleSubstepBegin(xrt_ctx* ctx, int inst);  // after predict: hash when due
int ParticleIter(xrt_ctx* ctx, int inst);          // iteration start: collide + apply
int ParticleSubstepEnd(xrt_ctx* ctx, int inst);    // after finalize
int ParticleStateReset(xrt_ctx* ctx, int inst);    // xrt_set_state: neighbor lists rebuilt
// binding_api.hip: nearest-face records, binding generation and skinning set-up at
// xrt_start (after ContactStart); the binding solve at the end of every iteration (an
// instance, or -1 for every instance)
bool BindingsOn(const xrt_ctx* ctx);
int BindingStart(xrt_ctx* ctx);
int BindingIter(xrt_ctx* ctx, int inst);
// fem_api.hip: the quality mode's FEM constraints.  FemPrepare builds them when the
// constraint model is FEM (xrt_start, xrt_set_params); FemIter runs one iteration's strain
// and isometric bending colors of an instance, or of every instance (-1)
bool FemOn(const xrt_ctx* ctx);
FemConsts FemConstsOf(const xrt_params& p);  // at the substep dt
int FemPrepare(xrt_ctx* ctx);
int FemIter(xrt_ctx* ctx, int inst, int it);
// adjoint_api.hip: one substep of every instance as wide launches with the contact phases
// above; with a tape owner and sub >= 0 the operators' inputs go to tape slot `sub`
int WideSubstep(xrt_ctx* ctx, const StepConsts& k, const BendMode& bm, AdjointData* tape, int sub);
}  // namespace xrt
