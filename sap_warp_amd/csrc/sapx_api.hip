// C ABI of libsapx (declared in include/sapx.h): model upload + validation,
// per-world topology template, solver / pipeline objects and launches.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/sapx.h"
#include "sapx_device.h"

#include "sapx_params.h"

namespace sapx {
hipError_t launch_step(const StepParams& P, hipStream_t stream);
hipError_t launch_wave_np16(const StepParams& P, int cb, hipStream_t stream);
hipError_t launch_wave_np32(const StepParams& P, int cb, hipStream_t stream);
hipError_t launch_wave_np48(const StepParams& P, int cb, hipStream_t stream);
hipError_t launch_wave_np64(const StepParams& P, int cb, hipStream_t stream);
hipError_t launch_collide(const CollideParams& P, const PlaceParams& Q, hipStream_t stream);
hipError_t collide_static_scratch(int** out);
hipError_t launch_batch_16(const BatchMap& B, int blocks, size_t smem, hipStream_t stream);
hipError_t launch_batch_32(const BatchMap& B, int blocks, size_t smem, hipStream_t stream);
hipError_t launch_batch_64(const BatchMap& B, int blocks, size_t smem, hipStream_t stream);
size_t collide_lds_bytes(const DevModel& m, int C);
hipError_t launch_world_order(const int* cworld_start, const int* iters, int W, int C, int* order, hipStream_t stream);
hipError_t launch_features(const FeatureParams& P, hipStream_t stream);
hipError_t launch_group_contacts(const GroupParams& G, hipStream_t stream);
void set_last_error(const char* msg);
}  // namespace sapx

using namespace sapx;

namespace {

thread_local std::string g_last_error = "";
// device histogram: iteration bins 0..255, line searches, world-steps, learned starts taken / rejected
constexpr int kHistSize = 260;
}  // namespace

// error text of the other translation units (sapx_mlp.hip)
void sapx::set_last_error(const char* msg) { g_last_error = msg; }

namespace {

int32_t fail(int32_t code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) return fail(SAPX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int32_t upload(std::vector<void*>& owned, const T* host, size_t count, const T** dev_out) {
    if (count == 0) {
        *dev_out = nullptr;
        return SAPX_OK;
    }
    void* d = nullptr;
    HIP_TRY(hipMalloc(&d, count * sizeof(T)));
    owned.push_back(d);
    HIP_TRY(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    *dev_out = reinterpret_cast<const T*>(d);
    return SAPX_OK;
}

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

struct SapxModel {
    DevModel dm;
    std::vector<void*> owned;
    std::vector<int> h_shape_world, h_shape_type;  // host copies for pipeline set-up
    std::vector<float> h_shape_scale;
    int32_t world_count = 0, joint_count = 0, dof_count = 0, coord_count = 0, body_count = 0, shape_count = 0;
    ~SapxModel() {
        for (void* p : owned) (void)hipFree(p);
    }
};

// device tables of a model, for the single-large-world unit
const sapx::DevModel& sapx::model_dev(const SapxModel* m) { return m->dm; }

struct SapxSolver {
    const SapxModel* model = nullptr;
    SapxSolverOpts opts{};
    double* v_guess = nullptr;
    int use_guess = 0;
    int* st = nullptr;  // 4 x W
    unsigned long long* hist = nullptr;  // kHistSize accumulators
    double* dbg = nullptr;
    size_t dbg_off[9] = {0};
    size_t dbg_size[9] = {0};
    int npad = 0, js = 0, ast = 0;
    int lds[17] = {0};
    int smem_bytes = 0;
    // large worlds: A, H and J of every world in device memory (fused kernel), floats
    float* big_ws = nullptr;
    size_t big_stride = 0, big_A = 0, big_H = 0, big_J = 0;
    // one-wave-per-world kernel: LDS layouts of the contact variants (16 / 32 / 64 slots; a
    // variant whose layout does not fit has bytes == 0) and the current schedule
    bool wave_ok = false;
    WaveLayout wl16{}, wl32{}, wl64{};
    SapxSchedule sched{};
    // the 64-slot variant runs on a second stream (default priority), forked from and joined
    // back into the caller's stream, so its few long worlds overlap the 16-slot launch
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // most-expensive-first dispatch order of the worlds (wave kernels), rebuilt every step
    int* order = nullptr;
    // learned warm start: feature layout and the caller's correction / target buffers
    int feat_count = 0;
    int frame_qs = -1;
    int* feat_off = nullptr;  // device, J+1
    const float* learn_delta = nullptr;
    float* learn_target = nullptr;
    // grouping buffers for ungrouped contact lists
    int group_capacity = 0;
    int* g_keys = nullptr;
    int* g_keys_out = nullptr;
    int* g_vals = nullptr;
    int* g_vals_out = nullptr;
    int* g_world_start = nullptr;
    int* g_shape0 = nullptr;
    int* g_shape1 = nullptr;
    float* g_p0 = nullptr;
    float* g_p1 = nullptr;
    float* g_n = nullptr;
    float* g_m0 = nullptr;
    float* g_m1 = nullptr;
    void* cub_tmp = nullptr;
    size_t cub_tmp_bytes = 0;
    ~SapxSolver() {
        void* ptrs[] = {v_guess, st, hist, dbg, g_keys, g_keys_out, g_vals, g_vals_out, g_world_start, g_shape0, g_shape1,
                        g_p0, g_p1, g_n, g_m0, g_m1, cub_tmp, order, feat_off, big_ws};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (aux) (void)hipStreamDestroy(aux);
    }
};

struct SapxPipeline {
    const SapxModel* model = nullptr;
    int C = 0;
    int* st_i = nullptr;  // W x 2C staged contact shape ids
    float* st_f = nullptr;  // W x 11C staged contact margins and points
    int* wcount = nullptr;  // W staged counts
    int* unsupported = nullptr;
    // all-pairs broadphase: the pipeline's own candidate template (P < 0: the model's explicit pairs)
    int* t_pairs = nullptr;
    int P = -1;
    int has_convex = 0;
    ~SapxPipeline() {
        void* ptrs[] = {st_i, st_f, wcount, unsupported, t_pairs};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
};

// shape-type pairs (sorted, ta <= tb) the collide kernel handles: the analytic primitives of
// narrow_phase.py:304-567 and the MPR / GJK path for the other support-mapped pairs; a plane
// reaches the MPR / GJK path only when it is infinite (zero half extents, core.py:680-697)
static bool pair_supported(int ta, int tb, bool plane_infinite) {
    if (pair_primitive(ta, tb)) return true;
    if (!pair_convex(ta, tb)) return false;
    return ta != kGeoPlane || plane_infinite;
}
static bool plane_infinite(const float* scale) { return scale[0] == 0.0f && scale[1] == 0.0f; }

// Several solvers (world kinds) stepped by one launch per contact variant.  Each bound slot holds
// the kernel parameters of every group for one set of state / control / contact views.
struct SapxBatch {
    std::vector<SapxSolver*> s;
    StepParams* dparams = nullptr;  // device: [slot 0..1][variant small, 64][group]
    std::vector<StepParams> hparams;
    struct Views {
        std::vector<SapxStateView> in, out;
        std::vector<SapxControlView> ctl;
        std::vector<SapxContactsView> c;
        double dt = 0.0;
        int bound = 0;
        std::vector<int> use_guess;  // each solver's warm-start flag when the slot was uploaded
    } views[2];
    int small = 16;
    int fork = 0;
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    ~SapxBatch() {
        if (dparams) (void)hipFree(dparams);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (aux) (void)hipStreamDestroy(aux);
    }
};

extern "C" {

int32_t sapx_version(void) { return 100; }

const char* sapx_last_error(void) { return g_last_error.c_str(); }

int32_t sapx_solver_feature_count(const SapxSolver* s) { return s ? s->feat_count : 0; }

int32_t sapx_features(SapxSolver* s, const SapxStateView* in, const SapxControlView* ctl, const SapxContactsView* c,
                      const float* prev_target, float* features, void* stream) {
    if (!s || !in || !features) return fail(SAPX_ERR_INVALID, "null argument");
    if (!in->joint_q || !in->joint_qd || !in->body_q) return fail(SAPX_ERR_INVALID, "state view needs joint_q, joint_qd and body_q");
    const DevModel& dm = s->model->dm;
    if (dm.NB > 64) return fail(SAPX_ERR_UNSUPPORTED, "features support at most 64 bodies per world");
    FeatureParams F;
    std::memset(&F, 0, sizeof(F));
    F.m = dm;
    F.jq = in->joint_q;
    F.jqd = in->joint_qd;
    F.bq = in->body_q;
    if (ctl) {
        F.tpos = ctl->joint_target_pos;
        F.tvel = ctl->joint_target_vel;
    }
    if (c && c->count) {
        if (!c->world_start) return fail(SAPX_ERR_UNSUPPORTED, "features need a world-grouped contact list (world_start)");
        F.cworld_start = c->world_start;
        F.cshape0 = c->shape0; F.cshape1 = c->shape1;
        F.cpoint0 = c->point0; F.cpoint1 = c->point1; F.cnormal = c->normal;
        F.cmargin0 = c->margin0; F.cmargin1 = c->margin1;
    }
    F.C = s->opts.max_rigid_contact;
    F.feat_off = s->feat_off;
    F.frame_qs = s->frame_qs;
    F.prev_target = prev_target;
    F.F = s->feat_count;
    F.out = features;
    HIP_TRY(launch_features(F, reinterpret_cast<hipStream_t>(stream)));
    return SAPX_OK;
}

int32_t sapx_solver_set_warm_start_io(SapxSolver* s, const float* start_correction, float* target_out) {
    if (!s) return fail(SAPX_ERR_INVALID, "null argument");
    if ((start_correction || target_out) && s->sched.kernel != SAPX_KERNEL_WAVE)
        return fail(SAPX_ERR_UNSUPPORTED, "the learned warm start needs the one-wave step kernel");
    s->learn_delta = start_correction;
    s->learn_target = target_out;
    return SAPX_OK;
}

int32_t sapx_solver_get_schedule(const SapxSolver* s, SapxSchedule* out) {
    if (!s || !out) return fail(SAPX_ERR_INVALID, "null argument");
    *out = s->sched;
    return SAPX_OK;
}

int32_t sapx_solver_set_schedule(SapxSolver* s, const SapxSchedule* in) {
    if (!s || !in) return fail(SAPX_ERR_INVALID, "null argument");
    if (in->kernel != SAPX_KERNEL_WAVE && in->kernel != SAPX_KERNEL_FUSED) return fail(SAPX_ERR_INVALID, "unknown kernel");
    if (in->kernel == SAPX_KERNEL_WAVE) {
        if (!s->wave_ok) return fail(SAPX_ERR_UNSUPPORTED, "the model does not fit the one-wave step kernel");
        const int c = in->small_capacity;
        if (c != 16 && c != 32 && c != 64) return fail(SAPX_ERR_INVALID, "small_capacity must be 16, 32 or 64");
        const WaveLayout& L = c == 16 ? s->wl16 : (c == 32 ? s->wl32 : s->wl64);
        if (L.bytes == 0) return fail(SAPX_ERR_UNSUPPORTED, "that contact variant does not fit in LDS");
        if (s->opts.max_rigid_contact > c && s->wl64.bytes == 0)
            return fail(SAPX_ERR_UNSUPPORTED, "the 64-slot variant does not fit in LDS");
    } else if (s->learn_delta || s->learn_target) {
        return fail(SAPX_ERR_UNSUPPORTED, "the learned warm start needs the one-wave step kernel");
    } else if (s->opts.line_search == SAPX_LS_MONOTONE_DECAY) {
        return fail(SAPX_ERR_UNSUPPORTED, "monotone_decay runs in the one-wave step kernel only");
    }
    SapxSchedule t = *in;
    t.fork = t.fork ? 1 : 0;
    t.world_order = t.world_order ? 1 : 0;
    if (t.fork && !s->aux) t.fork = 0;  // capacity <= 16: a single variant, nothing to fork
    s->sched = t;
    return SAPX_OK;
}

int32_t sapx_model_create(const SapxModelDesc* d, SapxModel** out) {
    if (!d || !out) return fail(SAPX_ERR_INVALID, "null argument");
    *out = nullptr;
    const int W = d->world_count;
    if (W <= 0) return fail(SAPX_ERR_INVALID, "world_count must be positive");
    if (d->joint_count <= 0 || d->dof_count <= 0)
        return fail(SAPX_ERR_INVALID, "SolverSAP requires a model with articulated joint DOFs.");
    if (d->joint_count % W || d->dof_count % W || d->coord_count % W || d->body_count % W)
        return fail(SAPX_ERR_INVALID, "joints/dofs/coords/bodies must divide evenly across worlds");
    const int J = d->joint_count / W, N = d->dof_count / W, NQ = d->coord_count / W, NB = d->body_count / W;
    // a world above the step kernels' limits (64 joints / bodies, 256 dofs) is stepped alone, over
    // the whole GPU (sapx_bigworld); sapx_solver_create rejects it
    const bool large = J > kMaxTemplate || NB > kMaxTemplate || N > kMaxDofFused;
    if (large && W != 1)
        return fail(SAPX_ERR_CAPACITY, "more than 64 joints or bodies or 256 dofs per world: one world only");
    // ---- topology template + equal-world validation
    std::vector<int> t_jtype(J), t_jparent(J), t_jchild(J), t_jqs(J + 1), t_jds(J + 1), t_jlin(J, 0);
    for (int w = 0; w < W; ++w) {
        for (int j = 0; j < J; ++j) {
            const int g = w * J + j;
            const int jt = d->joint_type[g];
            const int par = d->joint_parent[g] >= 0 ? d->joint_parent[g] - w * NB : -1;
            const int chi = d->joint_child[g] - w * NB;
            const int qs = d->joint_q_start[g] - w * NQ;
            const int ds = d->joint_qd_start[g] - w * N;
            if (w == 0) {
                t_jtype[j] = jt; t_jparent[j] = par; t_jchild[j] = chi; t_jqs[j] = qs; t_jds[j] = ds;
                if (jt == kJointD6) {
                    // D6: linear then angular axes, one coordinate per dof (sap_helpers.py:1164-1243)
                    if (!d->joint_dof_dim) return fail(SAPX_ERR_INVALID, "D6 joints need joint_dof_dim");
                    const int nl = d->joint_dof_dim[2 * g], na = d->joint_dof_dim[2 * g + 1];
                    const int nd = d->joint_qd_start[g + 1] - d->joint_qd_start[g];
                    const int nq = d->joint_q_start[g + 1] - d->joint_q_start[g];
                    if (nl < 0 || nl > 3 || na < 0 || na > 3 || nl + na != nd || nq != nd)
                        return fail(SAPX_ERR_INVALID, "D6 joint " + std::to_string(j) +
                                    ": needs 0-3 linear and 0-3 angular axes, one coordinate per dof");
                    if (nd == 0) return fail(SAPX_ERR_UNSUPPORTED, "D6 joint without axes (use a fixed joint)");
                    t_jlin[j] = nl;
                }
                if (chi < 0 || chi >= NB) return fail(SAPX_ERR_INVALID, "joint child outside its world");
                if (par >= NB) return fail(SAPX_ERR_INVALID, "joint parent outside its world");
            } else if (jt != t_jtype[j] || par != t_jparent[j] || chi != t_jchild[j] || qs != t_jqs[j] || ds != t_jds[j] ||
                       (jt == kJointD6 && d->joint_dof_dim[2 * g] != t_jlin[j])) {
                return fail(SAPX_ERR_UNSUPPORTED, "worlds must share one topology (world " + std::to_string(w) + ")");
            }
        }
    }
    t_jqs[J] = NQ;
    t_jds[J] = N;
    // body -> joint, depth / levels, ancestor + subtree masks
    std::vector<int> body_joint(NB, -1);
    for (int j = 0; j < J; ++j) {
        if (body_joint[t_jchild[j]] != -1) return fail(SAPX_ERR_INVALID, "body is the child of two joints");
        body_joint[t_jchild[j]] = j;
    }
    std::vector<int> depth(J, 0);
    for (int j = 0; j < J; ++j) {
        const int p = t_jparent[j];
        if (p >= 0) {
            const int pj = body_joint[p];
            if (pj < 0) return fail(SAPX_ERR_UNSUPPORTED, "parent body without a joint");
            if (pj >= j) return fail(SAPX_ERR_INVALID, "joints must be ordered parents-first");
            depth[j] = depth[pj] + 1;
        }
    }
    int nlev = 0;
    for (int j = 0; j < J; ++j) nlev = std::max(nlev, depth[j] + 1);
    std::vector<int> level_start(nlev + 1, 0), level_joint;
    for (int L = 0; L < nlev; ++L) {
        level_start[L] = (int)level_joint.size();
        for (int j = 0; j < J; ++j)
            if (depth[j] == L) level_joint.push_back(j);
    }
    level_start[nlev] = (int)level_joint.size();
    std::vector<unsigned long long> anc(NB, 0ull);
    for (int b = 0; b < NB; ++b) {
        int j = body_joint[b];
        while (j >= 0) {
            if (j < 64) anc[b] |= 1ull << j;  // masks of the step kernels (64 joints at most)
            const int p = t_jparent[j];
            j = p >= 0 ? body_joint[p] : -1;
        }
    }
    // child-body lists (CSR) for the deterministic reverse-level composite sweep
    std::vector<int> child_start(NB + 1, 0), child_list;
    for (int c = 0; c < NB; ++c) {
        child_start[c] = (int)child_list.size();
        for (int j = 0; j < J; ++j)
            if (t_jparent[j] == c) child_list.push_back(t_jchild[j]);
    }
    child_start[NB] = (int)child_list.size();
    if (child_list.empty()) child_list.push_back(0);
    // dof maps (contact_solve.py:3681-3720) and articulation dof ranges (3722-3760)
    std::vector<int> dof_joint(N), dof_coord(N, -1), dof_target(N), dof_limit(N, 0);
    for (int j = 0; j < J; ++j) {
        const int nd = t_jds[j + 1] - t_jds[j];
        for (int a = 0; a < nd; ++a) {
            const int dd = t_jds[j] + a;
            dof_joint[dd] = j;
            dof_target[dd] = dd;
        }
        const int jt = t_jtype[j];
        if (jt == kJointFree || jt == kJointDistance) {
            for (int a = 0; a < 3; ++a) {
                dof_target[t_jds[j] + a] = t_jds[j] + a + 3;
                dof_target[t_jds[j] + a + 3] = t_jds[j] + a;
            }
        }
        if (jt == kJointPrismatic || jt == kJointRevolute || jt == kJointD6) {
            for (int a = 0; a < nd; ++a) {
                dof_coord[t_jds[j] + a] = t_jqs[j] + a;
                if (nd == 1) dof_limit[t_jds[j] + a] = 1;
            }
        }
    }
    std::vector<int> art_start(NB, -1), art_count(NB, 0);
    for (int a = 0; a < d->articulation_count; ++a) {
        const int js0 = d->articulation_start[a], js1 = d->articulation_start[a + 1];
        if (js0 >= J) break;  // articulations of world 0 only
        int st = 1 << 30, cnt = 0;
        for (int j = js0; j < js1 && j < J; ++j) {
            const int nd = t_jds[j + 1] - t_jds[j];
            if (nd > 0) { st = std::min(st, t_jds[j]); cnt += nd; }
        }
        if (cnt <= 0) continue;
        for (int j = js0; j < js1 && j < J; ++j) {
            art_start[t_jchild[j]] = st;
            art_count[t_jchild[j]] = cnt;
        }
    }
    // shapes: world blocks then globals
    const int Stot = d->shape_count;
    int S = 0;
    int first_global = Stot;
    for (int s = 0; s < Stot; ++s) {
        if (d->shape_world[s] < 0) { first_global = std::min(first_global, s); }
        else if (d->shape_world[s] == 0) ++S;
    }
    for (int s = first_global; s < Stot; ++s)
        if (d->shape_world[s] >= 0) return fail(SAPX_ERR_UNSUPPORTED, "global shapes must follow all world shapes");
    if (first_global != W * S) return fail(SAPX_ERR_UNSUPPORTED, "every world must own the same number of shapes");
    for (int s = 0; s < W * S; ++s)
        if (d->shape_world[s] != s / S) return fail(SAPX_ERR_UNSUPPORTED, "shapes must be grouped by world");
    for (int s = 0; s < Stot; ++s) {
        const int t = d->shape_type[s];
        if (t != kGeoPlane && !support_type(t))
            return fail(SAPX_ERR_UNSUPPORTED, "shape type " + std::to_string(t) + " is not in the native narrowphase");
    }
    // pair template: pairs grouped by world (stable), encoded local / global
    std::vector<std::vector<int>> per_world(W);
    for (int p = 0; p < d->pair_count; ++p) {
        const int a = d->shape_pairs[2 * p], b = d->shape_pairs[2 * p + 1];
        if (a < 0 || b < 0 || a >= Stot || b >= Stot) return fail(SAPX_ERR_INVALID, "pair references a bad shape");
        const int wa = d->shape_world[a], wb = d->shape_world[b];
        const int wp = wa >= 0 ? wa : wb;
        if (wp < 0) continue;  // global-global pairs never collide in this path
        if (wa >= 0 && wb >= 0 && wa != wb) continue;
        const int ea = wa >= 0 ? a - wa * S : -a - 1;
        const int eb = wb >= 0 ? b - wb * S : -b - 1;
        per_world[wp].push_back(ea);
        per_world[wp].push_back(eb);
    }
    for (int w = 1; w < W; ++w)
        if (per_world[w] != per_world[0])
            return fail(SAPX_ERR_UNSUPPORTED, "candidate pairs must repeat identically in every world");
    // X_c identity flags
    std::vector<int> xcid(d->joint_count);
    for (int g = 0; g < d->joint_count; ++g) {
        const float* x = d->joint_X_c + 7 * g;
        xcid[g] = (x[0] == 0.0f && x[1] == 0.0f && x[2] == 0.0f && x[3] == 0.0f && x[4] == 0.0f && x[5] == 0.0f &&
                   x[6] == 1.0f);
    }
    SapxModel* m = new SapxModel();
    DevModel& dm = m->dm;
    m->h_shape_world.assign(d->shape_world, d->shape_world + Stot);
    m->h_shape_type.assign(d->shape_type, d->shape_type + Stot);
    m->h_shape_scale.assign(d->shape_scale, d->shape_scale + 3 * (size_t)Stot);
    dm.W = W; dm.J = J; dm.N = N; dm.NQ = NQ; dm.NB = NB; dm.S = S;
    dm.P = (int)per_world[0].size() / 2;
    dm.num_levels = nlev;
    dm.total_shapes = Stot;
    dm.global_shape_base = W * S;
    dm.has_convex = 0;
    for (size_t k = 0; k + 1 < per_world[0].size(); k += 2) {
        const int ea = per_world[0][k], eb = per_world[0][k + 1];
        int sa = ea >= 0 ? ea : -ea - 1, sb = eb >= 0 ? eb : -eb - 1;
        int ta = d->shape_type[sa], tb = d->shape_type[sb];
        if (ta > tb) { std::swap(ta, tb); std::swap(sa, sb); }
        if (!pair_supported(ta, tb, plane_infinite(d->shape_scale + 3 * (size_t)sa))) {
            delete m;
            return fail(SAPX_ERR_UNSUPPORTED, "candidate pair of shape types " + std::to_string(ta) + " and " +
                                                  std::to_string(tb) + " is not in the native narrowphase");
        }
        if (pair_convex(ta, tb)) dm.has_convex = 1;
    }
    int32_t rc = SAPX_OK;
    auto& o = m->owned;
#define UP(dst, src, cnt)                                   \
    if (rc == SAPX_OK) rc = upload(o, src, (size_t)(cnt), &dm.dst);
    UP(t_jtype, t_jtype.data(), J);
    UP(t_jparent, t_jparent.data(), J);
    UP(t_jchild, t_jchild.data(), J);
    UP(t_jqs, t_jqs.data(), J + 1);
    UP(t_jds, t_jds.data(), J + 1);
    UP(t_level_start, level_start.data(), nlev + 1);
    UP(t_level_joint, level_joint.data(), J);
    UP(t_dof_joint, dof_joint.data(), N);
    UP(t_dof_coord, dof_coord.data(), N);
    UP(t_dof_target, dof_target.data(), N);
    UP(t_dof_limit_ok, dof_limit.data(), N);
    UP(t_body_joint, body_joint.data(), NB);
    UP(t_body_anc, anc.data(), NB);
    UP(t_body_art_start, art_start.data(), NB);
    UP(t_child_start, child_start.data(), NB + 1);
    UP(t_child, child_list.data(), child_list.size());
    {
        // packed topology template (one coalesced copy into LDS per world)
        std::vector<int> topo;
        auto add = [&](int field, const int* src, size_t cnt) {
            dm.topo_off[field] = (int)topo.size();
            topo.insert(topo.end(), src, src + cnt);
        };
        add(TP_JTYPE, t_jtype.data(), J);
        add(TP_JPARENT, t_jparent.data(), J);
        add(TP_JCHILD, t_jchild.data(), J);
        add(TP_JQS, t_jqs.data(), J + 1);
        add(TP_JDS, t_jds.data(), J + 1);
        add(TP_LEVEL_START, level_start.data(), nlev + 1);
        add(TP_LEVEL_JOINT, level_joint.data(), J);
        add(TP_DOF_JOINT, dof_joint.data(), N);
        add(TP_DOF_COORD, dof_coord.data(), N);
        add(TP_DOF_TARGET, dof_target.data(), N);
        add(TP_DOF_LIMIT_OK, dof_limit.data(), N);
        add(TP_BODY_ART_START, art_start.data(), NB);
        add(TP_BODY_ART_COUNT, art_count.data(), NB);
        add(TP_CHILD_START, child_start.data(), NB + 1);
        add(TP_CHILD, child_list.data(), child_list.size());
        {
            // root body of each body's tree: the dynamics of a tree are written about its root's
            // origin, so a forest (several free bodies, several fixed-base trees) keeps every
            // body's spatial inertia about a nearby point
            std::vector<int> root(NB);
            for (int b = 0; b < NB; ++b) {
                int r = b;
                for (int guard = 0; guard <= NB && body_joint[r] >= 0 && t_jparent[body_joint[r]] >= 0; ++guard)
                    r = t_jparent[body_joint[r]];
                root[b] = r;
            }
            add(TP_BODY_ROOT, root.data(), NB);
        }
        add(TP_JLIN, t_jlin.data(), J);
        if (topo.size() % 2) topo.push_back(0);  // 8-byte alignment of the u64 masks
        auto add64 = [&](int field, const std::vector<unsigned long long>& v) {
            dm.topo_off[field] = (int)topo.size();
            for (unsigned long long x : v) {
                topo.push_back((int)(x & 0xffffffffull));
                topo.push_back((int)(x >> 32));
            }
        };
        // dof-level ancestry: bit d = dof d's joint lies on the root path
        std::vector<unsigned long long> anc_dof(NB, 0ull), dof_anc(N, 0ull);
        for (int b = 0; b < NB; ++b)
            for (int d = 0; d < N && d < 64; ++d)  // read only by the one-wave kernel (n <= 64)
                if (dof_joint[d] < 64 && ((anc[b] >> dof_joint[d]) & 1ull)) anc_dof[b] |= 1ull << d;
        for (int d = 0; d < N; ++d) dof_anc[d] = anc_dof[t_jchild[dof_joint[d]]];
        add64(TP_BODY_ANC, anc);
        add64(TP_BODY_ANC_DOF, anc_dof);
        add64(TP_DOF_ANC, dof_anc);
        {
            // ancestor tables for pointer jumping: round r links each body to the body 2^r
            // joints up its root path
            std::vector<int> cur(NB, -1), ups;
            for (int b = 0; b < NB; ++b) cur[b] = body_joint[b] >= 0 ? t_jparent[body_joint[b]] : -1;
            int rounds = 0;
            while (std::any_of(cur.begin(), cur.end(), [](int u) { return u >= 0; })) {
                ups.insert(ups.end(), cur.begin(), cur.end());
                std::vector<int> nxt(NB, -1);
                for (int b = 0; b < NB; ++b) nxt[b] = cur[b] >= 0 ? cur[cur[b]] : -1;
                cur.swap(nxt);
                ++rounds;
            }
            dm.fk_rounds = rounds;
            add(TP_BODY_UP, ups.data(), ups.size());
        }
        while (topo.size() % 4) topo.push_back(0);  // 16-byte copies into LDS
        dm.topo_ints = (int)topo.size();
        UP(topo, topo.data(), topo.size());
    }
    {
        // per-world parameter blocks in the one-wave kernel's LDS layout (one contiguous copy)
        const int wpf = 16 * J + 4 * N + 14 * NB;
        dm.wp_stride = (wpf + 3) / 4 * 4;
        std::vector<float> wp((size_t)W * dm.wp_stride, 0.0f);
        for (int w = 0; w < W; ++w) {
            float* o = wp.data() + (size_t)w * dm.wp_stride;
            for (int j = 0; j < J; ++j)
                for (int k = 0; k < 7; ++k) {
                    o[j * 8 + k] = d->joint_X_p[((size_t)w * J + j) * 7 + k];
                    o[J * 8 + j * 8 + k] = d->joint_X_c[((size_t)w * J + j) * 7 + k];
                }
            float* ax = o + 16 * J;
            for (int i = 0; i < N; ++i)
                for (int k = 0; k < 3; ++k) ax[i * 4 + k] = d->joint_axis[((size_t)w * N + i) * 3 + k];
            float* com = ax + 4 * N;
            float* in = com + 4 * NB;
            float* ms = in + 9 * NB;
            for (int b = 0; b < NB; ++b) {
                for (int k = 0; k < 3; ++k) com[b * 4 + k] = d->body_com[((size_t)w * NB + b) * 3 + k];
                for (int k = 0; k < 9; ++k) in[b * 9 + k] = d->body_inertia[((size_t)w * NB + b) * 9 + k];
                ms[b] = d->body_mass[(size_t)w * NB + b];
            }
        }
        UP(world_params, wp.data(), wp.size());
    }
    UP(t_body_art_count, art_count.data(), NB);
    UP(t_pairs, per_world[0].data(), per_world[0].size());
    UP(j_xc_identity, xcid.data(), d->joint_count);
    UP(joint_axis, d->joint_axis, (size_t)d->dof_count * 3);
    UP(joint_X_p, d->joint_X_p, (size_t)d->joint_count * 7);
    UP(joint_X_c, d->joint_X_c, (size_t)d->joint_count * 7);
    UP(armature, d->joint_armature, d->dof_count);
    UP(body_mass, d->body_mass, d->body_count);
    UP(body_inertia, d->body_inertia, (size_t)d->body_count * 9);
    UP(body_com, d->body_com, (size_t)d->body_count * 3);
    UP(gravity, d->gravity, (size_t)W * 3);
    UP(target_mode, d->joint_target_mode, d->dof_count);
    UP(target_ke, d->joint_target_ke, d->dof_count);
    UP(target_kd, d->joint_target_kd, d->dof_count);
    UP(effort_limit, d->joint_effort_limit, d->dof_count);
    UP(limit_lower, d->joint_limit_lower, d->dof_count);
    UP(limit_upper, d->joint_limit_upper, d->dof_count);
    UP(limit_ke, d->joint_limit_ke, d->dof_count);
    UP(limit_kd, d->joint_limit_kd, d->dof_count);
    UP(shape_body, d->shape_body, Stot);
    UP(shape_type, d->shape_type, Stot);
    UP(shape_transform, d->shape_transform, (size_t)Stot * 7);
    UP(shape_scale, d->shape_scale, (size_t)Stot * 3);
    UP(shape_margin, d->shape_margin, Stot);
    UP(shape_gap, d->shape_gap, Stot);
    UP(shape_radius, d->shape_collision_radius, Stot);
    UP(shape_mu, d->shape_material_mu, Stot);
    UP(shape_ke, d->shape_material_ke, Stot);
    UP(shape_tau, d->shape_material_tau, Stot);
#undef UP
    if (rc != SAPX_OK) {
        delete m;
        return rc;
    }
    m->world_count = W;
    m->joint_count = d->joint_count;
    m->dof_count = d->dof_count;
    m->coord_count = d->coord_count;
    m->body_count = d->body_count;
    m->shape_count = Stot;
    *out = m;
    return SAPX_OK;
}

void sapx_model_destroy(SapxModel* model) { delete model; }

int32_t sapx_solver_create(const SapxModel* model, const SapxSolverOpts* opts, SapxSolver** out) {
    if (!model || !opts || !out) return fail(SAPX_ERR_INVALID, "null argument");
    *out = nullptr;
    const DevModel& dm = model->dm;
    const int C = opts->max_rigid_contact;
    if (C <= 0) return fail(SAPX_ERR_INVALID, "max_rigid_contact must be positive");
    if (dm.J > kMaxTemplate || dm.NB > kMaxTemplate || dm.N > kMaxDofFused)
        return fail(SAPX_ERR_CAPACITY, "more than 64 joints or bodies or 256 dofs per world: step the world with "
                                       "sapx_bigworld");
    if (opts->line_search != SAPX_LS_ARMIJO_DECAY && opts->line_search != SAPX_LS_EXACT_ROOT &&
        opts->line_search != SAPX_LS_MONOTONE_DECAY)
        return fail(SAPX_ERR_UNSUPPORTED, "line search variant not in the native step");
    if (opts->line_search == SAPX_LS_ARMIJO_DECAY && !(opts->rho > 0.0 && opts->rho < 1.0))
        return fail(SAPX_ERR_INVALID, "rho must lie in (0, 1)");
    SapxSolver* s = new SapxSolver();
    s->model = model;
    s->opts = *opts;
    const int n = dm.N, nb = dm.NB;
    s->npad = (int)round_up(n, 16);
    // odd row strides keep row-per-thread LDS access free of bank conflicts
    s->js = (n % 2) ? n : n + 1;
    s->ast = s->npad + 1;
    const int npad = s->npad, hs = npad + 1, js = s->js;
    // LDS layout (see Smem in sapx_step.hip)
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = round_up(off + bytes, 16);
        return (int)o;
    };
    const size_t a_bytes = (size_t)n * s->ast * 4;                                                  // A
    const size_t h_bytes = std::max((size_t)npad * hs * 4, (size_t)C * 104);                        // H / contact scratch
    const size_t j_bytes = std::max((size_t)3 * C * js * 4, (size_t)(50 * nb + 14 * n + 2) * 4 + (size_t)dm.J * 64);  // J / kinematics
    // A, H and J stay in LDS while the whole world fits (with room for the rest); a world above 64
    // dofs, or one whose matrices alone pass 112 KiB, keeps them in a per-world block of device
    // memory (L2 / MALL resident while its workgroup runs) and the vectors in LDS
    const bool big = npad > 64 || a_bytes + h_bytes + j_bytes > 112 * 1024;
    if (big) {
        auto fl = [](size_t bytes) { return round_up(bytes, 64) / 4; };  // 64-byte aligned regions
        s->big_A = 0;
        s->big_H = fl(a_bytes);
        s->big_J = s->big_H + fl(h_bytes);
        s->big_stride = s->big_J + fl(j_bytes);
        s->lds[0] = s->lds[1] = s->lds[2] = 0;
    } else {
        s->lds[0] = take(a_bytes);
        s->lds[1] = take(h_bytes);
        s->lds[2] = take(j_bytes);
    }
    s->lds[3] = take((size_t)nb * 8 * 4);                                                // bq
    s->lds[4] = take((size_t)npad * 6 * 4);                                              // S
    s->lds[5] = take((size_t)C * 8 * 8);                                                 // ck (kCk = 8)
    s->lds[6] = take((size_t)C * 9 * 8);                                                 // cv
    s->lds[7] = take((size_t)C * 9 * 4);                                                 // cG
    s->lds[8] = take((size_t)C * 2 * 4);                                                 // cb
    s->lds[9] = take((size_t)12 * npad * 8);                                             // d64 (V_NUM rows)
    s->lds[10] = take((size_t)16 * npad * 4);                                            // d32 (>= F_NUM rows)
    s->lds[11] = take((size_t)4 * npad * 4);                                             // dflag
    s->lds[12] = take((size_t)8 * 8 * 8);                                                // red
    s->lds[13] = take((size_t)2 * dm.NQ * 4);                                            // q
    s->lds[14] = take((size_t)(dm.J * 16 + n * 4 + nb * 4 + nb * 9 + nb) * 4);             // per-world model params
    s->lds[15] = take((size_t)nb * 8 * 8);                                               // bqd (f64 poses)
    s->lds[16] = take((size_t)dm.topo_ints * 4);                                         // topology template
    s->smem_bytes = (int)off;
    if (s->smem_bytes > 160 * 1024) {
        delete s;
        return fail(SAPX_ERR_CAPACITY, "per-world LDS footprint exceeds 160 KiB; lower max_rigid_contact");
    }
    if (big) {
        const size_t bytes = (size_t)dm.W * s->big_stride * sizeof(float);
        if (hipMalloc(&s->big_ws, bytes) != hipSuccess) {
            delete s;
            return fail(SAPX_ERR_HIP, "hipMalloc failed for the large-world matrices");
        }
    }
    {
        // one wave per world: LDS layout per contact variant (see sapx_wave_impl.h)
        const int NP = npad;
        auto layout = [&](int CB) {
            WaveLayout L{};
            size_t o = 0;
            auto tk = [&](size_t bytes) {
                size_t r = o;
                o = round_up(o + bytes, 16);
                return (int)r;
            };
            L.js = NP + 1;  // zero-padded rows: the mat-vecs run over NP columns unguarded
            L.ls = NP + 1;
            L.o_topo = tk((size_t)dm.topo_ints * 4);
            L.o_mp = tk((size_t)dm.wp_stride * 4);
            L.o_q = tk((size_t)2 * dm.NQ * 4);
            L.o_bq = tk((size_t)nb * 8 * 4);
            L.o_bqd = tk((size_t)nb * 8 * 8);
            L.o_S = tk((size_t)n * 6 * 4);
            L.o_vec = tk((size_t)128 * 8);
            L.o_ck = tk((size_t)CB * 8 * 8);
            L.o_cG = tk((size_t)CB * 9 * 4);
            L.o_A = tk((size_t)n * L.ls * 4);
            // J rows then the factor rows; the kinematics scratch overlays both
            const size_t jbytes = round_up((size_t)3 * CB * L.js * 4, 16);
            const size_t ltbytes = ((size_t)NP * L.ls + 64) * 4;
            const size_t kin = (size_t)(54 * nb + 14 * n + 2) * 4 + (size_t)dm.J * 128 + (size_t)nb * 64;
            L.o_J = (int)o;
            L.o_Lt = (int)(o + jbytes);
            o = round_up(o + std::max(jbytes + ltbytes, kin), 16);
            L.bytes = (int)o;
            return L;
        };
        if (NP <= 64 && C <= 64) {
            auto fit = [&](int CB) {
                WaveLayout L = layout(CB);
                if (L.bytes > 160 * 1024) L.bytes = 0;
                return L;
            };
            s->wl16 = fit(16);
            s->wl32 = fit(32);
            s->wl64 = fit(64);
            // the 64-slot variant holds every world; the small ones need it beside them when C > slots
            s->wave_ok = s->wl64.bytes > 0 || (C <= 16 && s->wl16.bytes > 0) || (C <= 32 && s->wl32.bytes > 0);
        }
        // default schedule; SAPX_STEP_KERNEL ("fused", "wave64") and SAPX_WORLD_ORDER=0 override it
        // (comparisons and tests)
        const char* env = getenv("SAPX_STEP_KERNEL");
        const std::string kenv = env ? std::string(env) : std::string();
        const char* oenv = getenv("SAPX_WORLD_ORDER");
        s->sched.kernel = (s->wave_ok && kenv != "fused") ? SAPX_KERNEL_WAVE : SAPX_KERNEL_FUSED;
        s->sched.small_capacity = kenv == "wave64" ? 64 : 16;
        s->sched.fork = 1;
        s->sched.world_order = (oenv && std::string(oenv) == "0") ? 0 : 1;
        if (opts->line_search == SAPX_LS_MONOTONE_DECAY && s->sched.kernel != SAPX_KERNEL_WAVE) {
            delete s;
            return fail(SAPX_ERR_UNSUPPORTED, "monotone_decay runs in the one-wave step kernel only");
        }
        if (s->wave_ok && hipMalloc(&s->order, (size_t)dm.W * sizeof(int)) != hipSuccess) {
            delete s;
            return fail(SAPX_ERR_HIP, "hipMalloc failed for the world order");
        }
        if (s->wave_ok && C > 16) {
            if (hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
                delete s;
                return fail(SAPX_ERR_HIP, "stream / event creation failed for the solver");
            }
        }
    }
    const size_t W = dm.W, D = (size_t)dm.W * n;
    if (hipMalloc(&s->v_guess, D * sizeof(double)) != hipSuccess || hipMalloc(&s->st, 4 * W * sizeof(int)) != hipSuccess) {
        delete s;
        return fail(SAPX_ERR_HIP, "hipMalloc failed for solver buffers");
    }
    (void)hipMemset(s->v_guess, 0, D * sizeof(double));
    (void)hipMemset(s->st, 0, 4 * W * sizeof(int));
    if (hipMalloc(&s->hist, kHistSize * sizeof(unsigned long long)) != hipSuccess) {
        delete s;
        return fail(SAPX_ERR_HIP, "hipMalloc failed for the iteration histogram");
    }
    (void)hipMemset(s->hist, 0, kHistSize * sizeof(unsigned long long));
    {
        // feature layout of the learned warm start (joint blocks, then one block per body)
        const int J = dm.J;
        std::vector<int> jt(J), jqs(J + 1), jds(J + 1), off(J + 1);
        if (hipMemcpy(jt.data(), dm.t_jtype, (size_t)J * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(jqs.data(), dm.t_jqs, (size_t)(J + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(jds.data(), dm.t_jds, (size_t)(J + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            delete s;
            return fail(SAPX_ERR_HIP, "reading the topology template failed");
        }
        int o = 0;
        for (int j = 0; j < J; ++j) {
            off[j] = o;
            if (jt[j] == kJointFree || jt[j] == kJointDistance) {
                o += 10;
                if (s->frame_qs < 0) s->frame_qs = jqs[j];
            } else {
                o += 4 * (jds[j + 1] - jds[j]);
            }
        }
        off[J] = o;
        s->feat_count = o + 5 * dm.NB + dm.N;
        if (hipMalloc(&s->feat_off, (size_t)(J + 1) * 4) != hipSuccess ||
            hipMemcpy(s->feat_off, off.data(), (size_t)(J + 1) * 4, hipMemcpyHostToDevice) != hipSuccess) {
            delete s;
            return fail(SAPX_ERR_HIP, "hipMalloc failed for the feature layout");
        }
    }
    const size_t nd = (size_t)dm.N;
    const size_t sizes[9] = {D, D, D, W * C * 3, W * C, W * C, D, W * nd * nd, W * C * 3 * nd};
    size_t tot = 0;
    for (int i = 0; i < 9; ++i) {
        s->dbg_off[i] = tot;
        s->dbg_size[i] = sizes[i] * sizeof(double);
        tot += sizes[i];
    }
    if (opts->capture_snapshots) {
        if (hipMalloc(&s->dbg, tot * sizeof(double)) != hipSuccess) {
            delete s;
            return fail(SAPX_ERR_HIP, "hipMalloc failed for solver snapshot buffers");
        }
        (void)hipMemset(s->dbg, 0, tot * sizeof(double));
    }
    s->use_guess = 0;
    *out = s;
    return SAPX_OK;
}

void sapx_solver_destroy(SapxSolver* solver) { delete solver; }

int32_t sapx_solver_reset(SapxSolver* solver, void* stream) {
    if (!solver) return fail(SAPX_ERR_INVALID, "null solver");
    solver->use_guess = 0;
    (void)stream;
    return SAPX_OK;
}

static int32_t ensure_group_buffers(SapxSolver* s, int capacity) {
    if (capacity <= s->group_capacity) return SAPX_OK;
    void* ptrs[] = {s->g_keys, s->g_keys_out, s->g_vals, s->g_vals_out, s->g_shape0, s->g_shape1, s->g_p0, s->g_p1,
                    s->g_n, s->g_m0, s->g_m1, s->cub_tmp};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    const size_t c = (size_t)capacity;
    HIP_TRY(hipMalloc(&s->g_keys, c * 4));
    HIP_TRY(hipMalloc(&s->g_keys_out, c * 4));
    HIP_TRY(hipMalloc(&s->g_vals, c * 4));
    HIP_TRY(hipMalloc(&s->g_vals_out, c * 4));
    HIP_TRY(hipMalloc(&s->g_shape0, c * 4));
    HIP_TRY(hipMalloc(&s->g_shape1, c * 4));
    HIP_TRY(hipMalloc(&s->g_p0, c * 12));
    HIP_TRY(hipMalloc(&s->g_p1, c * 12));
    HIP_TRY(hipMalloc(&s->g_n, c * 12));
    HIP_TRY(hipMalloc(&s->g_m0, c * 4));
    HIP_TRY(hipMalloc(&s->g_m1, c * 4));
    if (!s->g_world_start) HIP_TRY(hipMalloc(&s->g_world_start, ((size_t)s->model->dm.W + 1) * 4));
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, s->g_keys, s->g_keys_out, s->g_vals, s->g_vals_out,
                                               capacity));
    size_t tmp2 = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, s->g_keys, s->g_world_start, s->model->dm.W + 1));
    s->cub_tmp_bytes = std::max(tmp, tmp2);
    HIP_TRY(hipMalloc(&s->cub_tmp, s->cub_tmp_bytes));
    s->group_capacity = capacity;
    return SAPX_OK;
}

// Kernel parameters of one step of solver s (all but the wave layout and the world order); an
// ungrouped contact list is grouped by world first (launches on st).
static int32_t fill_step_params(SapxSolver* s, const SapxStateView* in, const SapxStateView* out,
                                const SapxControlView* ctl, const SapxContactsView* c, double dt,
                                const double* dt_world, hipStream_t st, StepParams* out_params) {
    if (!s || !in || !out) return fail(SAPX_ERR_INVALID, "null argument");
    if (!in->joint_q || !in->joint_qd || !in->body_q || !out->joint_q || !out->joint_qd || !out->body_q)
        return fail(SAPX_ERR_INVALID, "state views need joint_q, joint_qd and body_q");
    if (!(dt > 0.0) && !dt_world) return fail(SAPX_ERR_INVALID, "dt must be positive");
    const DevModel& dm = s->model->dm;
    StepParams& P = *out_params;
    std::memset(&P, 0, sizeof(P));
    P.m = dm;
    P.jq_in = in->joint_q;
    P.jqd_in = in->joint_qd;
    P.bq_in = in->body_q;
    P.bf_in = in->body_f;
    if (ctl) {
        P.c_jf = ctl->joint_f;
        P.c_tpos = ctl->joint_target_pos;
        P.c_tvel = ctl->joint_target_vel;
        P.c_act = ctl->joint_act;
    }
    if (c && c->count) {
        if (c->world_start) {
            P.cworld_start = c->world_start;
            P.cshape0 = c->shape0; P.cshape1 = c->shape1;
            P.cpoint0 = c->point0; P.cpoint1 = c->point1; P.cnormal = c->normal;
            P.cmargin0 = c->margin0; P.cmargin1 = c->margin1;
        } else {
            // group an arbitrary flat list by world (stable radix sort on the world id)
            const int cap = c->capacity;
            if (cap > 0) {
                int32_t rc = ensure_group_buffers(s, cap);
                if (rc != SAPX_OK) return rc;
                GroupParams G;
                G.W = dm.W; G.NB = dm.NB; G.capacity = cap; G.count = c->count;
                G.shape_body = dm.shape_body;
                G.in_shape0 = c->shape0; G.in_shape1 = c->shape1;
                G.in_p0 = c->point0; G.in_p1 = c->point1; G.in_n = c->normal; G.in_m0 = c->margin0; G.in_m1 = c->margin1;
                G.keys = s->g_keys; G.vals = s->g_vals; G.keys_out = s->g_keys_out; G.vals_out = s->g_vals_out;
                G.world_start = s->g_world_start;
                G.out_shape0 = s->g_shape0; G.out_shape1 = s->g_shape1;
                G.out_p0 = s->g_p0; G.out_p1 = s->g_p1; G.out_n = s->g_n; G.out_m0 = s->g_m0; G.out_m1 = s->g_m1;
                G.cub_tmp = s->cub_tmp; G.cub_tmp_bytes = s->cub_tmp_bytes;
                HIP_TRY(launch_group_contacts(G, st));
                P.cworld_start = s->g_world_start;
                P.cshape0 = s->g_shape0; P.cshape1 = s->g_shape1;
                P.cpoint0 = s->g_p0; P.cpoint1 = s->g_p1; P.cnormal = s->g_n;
                P.cmargin0 = s->g_m0; P.cmargin1 = s->g_m1;
            }
        }
    }
    P.jq_out = out->joint_q;
    P.jqd_out = out->joint_qd;
    P.bq_out = out->body_q;
    P.bqd_out = out->body_qd;
    P.bf_out = out->body_f;
    P.v_guess = s->v_guess;
    P.use_guess = s->use_guess;
    const size_t W = dm.W;
    P.st_iters = s->st;
    P.st_ls = s->st + W;
    P.st_nc = s->st + 2 * W;
    P.st_flags = s->st + 3 * W;
    P.st_hist = s->hist;
    if (s->dbg) {
        P.dbg_v = s->dbg + s->dbg_off[0];
        P.dbg_vstar = s->dbg + s->dbg_off[1];
        P.dbg_v0 = s->dbg + s->dbg_off[2];
        P.dbg_gamma = s->dbg + s->dbg_off[3];
        P.dbg_phi0 = s->dbg + s->dbg_off[4];
        P.dbg_w = s->dbg + s->dbg_off[5];
        P.dbg_adiag = s->dbg + s->dbg_off[6];
        P.dbg_A = s->dbg + s->dbg_off[7];
        P.dbg_J = s->dbg + s->dbg_off[8];
    }
    P.dt = dt;
    P.dt_world = dt_world;
    P.C = s->opts.max_rigid_contact;
    const SapxSolverOpts& o = s->opts;
    P.fallback_mu = o.fallback_mu;
    P.fallback_k = o.fallback_stiffness;
    P.fallback_tau = o.fallback_tau_d;
    P.beta = o.contact_beta;
    P.sigma = o.contact_sigma;
    P.diag_shift = o.diag_shift;
    P.max_iter = o.max_iterations;
    P.opt_abs = o.optimality_abs_tol;
    P.opt_rel = o.optimality_rel_tol;
    P.cost_abs = o.cost_abs_tol;
    P.cost_rel = o.cost_rel_tol;
    P.ls_mode = o.line_search;
    P.ls_max_iter = o.line_search_max_iterations;
    P.armijo_c = o.armijo_c;
    P.rho = o.rho;
    P.rel_slop = o.line_search_relative_slop;
    P.weight_mode = o.contact_weight_mode;
    P.point_mode = o.contact_point_mode;
    P.integ_mode = o.position_integration;
    P.npad = s->npad;
    P.js = s->js;
    P.ast = s->ast;
    P.o_A = s->lds[0]; P.o_H = s->lds[1]; P.o_J = s->lds[2]; P.o_bq = s->lds[3]; P.o_S = s->lds[4];
    P.o_ck = s->lds[5]; P.o_cv = s->lds[6]; P.o_cG = s->lds[7]; P.o_cb = s->lds[8]; P.o_dv64 = s->lds[9];
    P.o_dv32 = s->lds[10]; P.o_dflag = s->lds[11]; P.o_red = s->lds[12]; P.o_q = s->lds[13]; P.o_misc = s->lds[14];
    P.o_bqd = s->lds[15];
    P.o_topo = s->lds[16];
    P.smem_bytes = s->smem_bytes;
    P.g_work = s->big_ws;
    P.g_stride = s->big_stride;
    P.g_A = s->big_A;
    P.g_H = s->big_H;
    P.g_J = s->big_J;
    return SAPX_OK;
}

int32_t sapx_step(SapxSolver* s, const SapxStateView* in, const SapxStateView* out, const SapxControlView* ctl,
                  const SapxContactsView* c, double dt, const double* dt_world, void* stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    StepParams P;
    {
        const int32_t rc = fill_step_params(s, in, out, ctl, c, dt, dt_world, st, &P);
        if (rc != SAPX_OK) return rc;
    }
    const DevModel& dm = s->model->dm;
    if (s->sched.kernel == SAPX_KERNEL_WAVE) {
        auto launch_on = [&](const StepParams& Q, int cb, hipStream_t sq) {
            switch (s->npad) {
                case 16: return launch_wave_np16(Q, cb, sq);
                case 32: return launch_wave_np32(Q, cb, sq);
                case 48: return launch_wave_np48(Q, cb, sq);
                default: return launch_wave_np64(Q, cb, sq);
            }
        };
        P.world_order = nullptr;
        P.learn_delta = s->learn_delta;
        P.learn_target = s->learn_target;
        if (s->order && s->sched.world_order) {
            HIP_TRY(launch_world_order(P.cworld_start, P.st_iters, (int)dm.W, P.C, s->order, st));
            P.world_order = s->order;
        }
        // contact variants: the small one holds every world when the capacity fits in it,
        // otherwise the 64-slot variant takes the worlds above its slots
        const int small = s->sched.small_capacity;
        const WaveLayout& wls = small == 16 ? s->wl16 : (small == 32 ? s->wl32 : s->wl64);
        if (P.C <= small || small == 64) {
            P.wl = wls;
            P.wave_min_nc = 0;
            HIP_TRY(launch_on(P, small, st));
        } else {
            StepParams Pb = P;
            Pb.wl = s->wl64;
            Pb.wave_min_nc = small + 1;
            P.wl = wls;
            P.wave_min_nc = 0;
            if (s->sched.fork) {
                // fork: the worlds above `small` contacts start first on the second stream, the
                // rest fill the remaining wave slots from the caller's stream; join before returning
                HIP_TRY(hipEventRecord(s->ev_fork, st));
                HIP_TRY(hipStreamWaitEvent(s->aux, s->ev_fork, 0));
                HIP_TRY(launch_on(Pb, 64, s->aux));
                HIP_TRY(hipEventRecord(s->ev_join, s->aux));
                HIP_TRY(launch_on(P, small, st));
                HIP_TRY(hipStreamWaitEvent(st, s->ev_join, 0));
            } else {
                HIP_TRY(launch_on(P, small, st));
                HIP_TRY(launch_on(Pb, 64, st));
            }
        }
    } else {
        if (s->learn_delta || s->learn_target)
            return fail(SAPX_ERR_UNSUPPORTED, "the learned warm start needs the one-wave step kernel");
        HIP_TRY(launch_step(P, st));
    }
    s->use_guess = 1;
    return SAPX_OK;
}

int32_t sapx_get_stats(SapxSolver* s, int32_t* newton_iterations, int32_t* line_search_iterations,
                       int32_t* contact_count, int32_t* flags, void* stream) {
    if (!s) return fail(SAPX_ERR_INVALID, "null solver");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const size_t W = s->model->dm.W;
    int32_t* dst[4] = {newton_iterations, line_search_iterations, contact_count, flags};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) HIP_TRY(hipMemcpyAsync(dst[k], s->st + k * W, W * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SAPX_OK;
}

int32_t sapx_get_histogram(SapxSolver* s, int64_t* hist, int32_t reset, void* stream) {
    if (!s || !hist) return fail(SAPX_ERR_INVALID, "null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(hist, s->hist, kHistSize * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (reset) HIP_TRY(hipMemsetAsync(s->hist, 0, kHistSize * sizeof(unsigned long long), st));
    return SAPX_OK;
}

int32_t sapx_solver_read(SapxSolver* s, int32_t which, void* dst, int64_t bytes, void* stream) {
    if (!s || !dst) return fail(SAPX_ERR_INVALID, "null argument");
    if (which < 0 || which > 8) return fail(SAPX_ERR_INVALID, "unknown buffer id");
    if (!s->dbg) return fail(SAPX_ERR_INVALID, "solver was created without capture_snapshots");
    if ((size_t)bytes != s->dbg_size[which]) return fail(SAPX_ERR_INVALID, "byte count does not match the buffer");
    HIP_TRY(hipMemcpyAsync(dst, s->dbg + s->dbg_off[which], (size_t)bytes, hipMemcpyDeviceToDevice,
                           reinterpret_cast<hipStream_t>(stream)));
    return SAPX_OK;
}

int32_t sapx_pipeline_create(const SapxModel* model, int32_t max_contacts_per_world, SapxPipeline** out) {
    if (!model || !out) return fail(SAPX_ERR_INVALID, "null argument");
    if (max_contacts_per_world <= 0) return fail(SAPX_ERR_INVALID, "max_contacts_per_world must be positive");
    *out = nullptr;
    SapxPipeline* p = new SapxPipeline();
    p->model = model;
    p->C = max_contacts_per_world;
    if (collide_lds_bytes(model->dm, p->C) > kCollideMaxLds) {
        delete p;
        return fail(SAPX_ERR_CAPACITY, "max_contacts_per_world too large for the collide kernel's LDS staging");
    }
    const size_t slots = (size_t)model->dm.W * p->C;
    bool ok = hipMalloc(&p->st_i, slots * 2 * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->st_f, slots * 11 * sizeof(float)) == hipSuccess &&
              hipMalloc(&p->wcount, (size_t)model->dm.W * sizeof(int)) == hipSuccess &&
              hipMalloc(&p->unsupported, sizeof(int)) == hipSuccess &&
              hipMemset(p->unsupported, 0, sizeof(int)) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        delete p;
        return fail(SAPX_ERR_HIP, "hipMalloc failed for pipeline buffers");
    }
    *out = p;
    return SAPX_OK;
}

void sapx_pipeline_destroy(SapxPipeline* pipeline) { delete pipeline; }

int32_t sapx_pipeline_create_all_pairs(const SapxModel* model, int32_t max_contacts_per_world,
                                       const int32_t* shape_collision_group, const int32_t* shape_flags,
                                       const int32_t* filter_pairs, int32_t filter_count, SapxPipeline** out) {
    if (!model || !out || !shape_collision_group) return fail(SAPX_ERR_INVALID, "null argument");
    if (filter_count < 0 || (filter_count > 0 && !filter_pairs)) return fail(SAPX_ERR_INVALID, "bad filter pairs");
    const DevModel& dm = model->dm;
    const int W = dm.W, S = dm.S, Stot = dm.total_shapes;
    const std::vector<int>& sw = model->h_shape_world;
    const std::vector<int>& st = model->h_shape_type;
    constexpr int kCollideShapes = 1 << 1;
    auto collides = [&](int s) { return !shape_flags || (shape_flags[s] & kCollideShapes) != 0; };
    // excluded pairs, canonical (min, max), sorted for binary search
    std::vector<std::pair<int, int>> excl;
    excl.reserve(filter_count);
    for (int k = 0; k < filter_count; ++k) {
        const int a = filter_pairs[2 * k], b = filter_pairs[2 * k + 1];
        excl.emplace_back(std::min(a, b), std::max(a, b));
    }
    std::sort(excl.begin(), excl.end());
    auto group_ok = [](int ga, int gb) {
        if (ga == 0 || gb == 0) return false;
        if (ga > 0) return ga == gb || gb < 0;
        return ga != gb;
    };
    std::vector<int> globals;
    for (int s = dm.global_shape_base; s < Stot; ++s)
        if (collides(s)) globals.push_back(s);
    // each world's slice (its colliding shapes, then the colliding global shapes) in lower-triangular
    // order; global-global pairs never collide in this path
    std::vector<int> tmpl;
    for (int w = 0; w < W; ++w) {
        std::vector<int> slice;
        for (int s = w * S; s < (w + 1) * S; ++s)
            if (collides(s)) slice.push_back(s);
        slice.insert(slice.end(), globals.begin(), globals.end());
        std::vector<int> mine;
        const int n = (int)slice.size();
        for (int r = 0; r < n; ++r)
            for (int c = r + 1; c < n; ++c) {
                const int a = std::min(slice[r], slice[c]), b = std::max(slice[r], slice[c]);
                const int wa = sw[a], wb = sw[b];
                if (wa < 0 && wb < 0) continue;
                if (wa >= 0 && wb >= 0 && wa != wb) continue;
                if (!group_ok(shape_collision_group[a], shape_collision_group[b])) continue;
                if (std::binary_search(excl.begin(), excl.end(), std::make_pair(a, b))) continue;
                mine.push_back(wa >= 0 ? a - wa * S : -a - 1);
                mine.push_back(wb >= 0 ? b - wb * S : -b - 1);
            }
        if (w == 0) tmpl = mine;
        else if (mine != tmpl)
            return fail(SAPX_ERR_UNSUPPORTED, "all-pairs candidates must repeat identically in every world");
    }
    int has_convex = 0;
    for (size_t k = 0; k + 1 < tmpl.size(); k += 2) {
        const int ea = tmpl[k], eb = tmpl[k + 1];
        int sa = ea >= 0 ? ea : -ea - 1, sb = eb >= 0 ? eb : -eb - 1;
        int ta = st[sa], tb = st[sb];
        if (ta > tb) { std::swap(ta, tb); std::swap(sa, sb); }
        if (pair_convex(ta, tb)) has_convex = 1;
        if (!pair_supported(ta, tb, plane_infinite(model->h_shape_scale.data() + 3 * (size_t)sa)))
            return fail(SAPX_ERR_UNSUPPORTED, "candidate pair of shape types " + std::to_string(ta) + " and " +
                                                  std::to_string(tb) + " is not in the native narrowphase");
    }
    int32_t rc = sapx_pipeline_create(model, max_contacts_per_world, out);
    if (rc != SAPX_OK) return rc;
    SapxPipeline* p = *out;
    p->P = (int)tmpl.size() / 2;
    p->has_convex = has_convex;
    DevModel dm_all = dm;
    dm_all.has_convex = has_convex;
    if (collide_lds_bytes(dm_all, p->C) > kCollideMaxLds) {
        delete p;
        *out = nullptr;
        return fail(SAPX_ERR_CAPACITY, "max_contacts_per_world too large for the collide kernel's LDS staging");
    }
    if (!tmpl.empty()) {
        if (hipMalloc(&p->t_pairs, tmpl.size() * sizeof(int)) != hipSuccess ||
            hipMemcpy(p->t_pairs, tmpl.data(), tmpl.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            delete p;
            *out = nullptr;
            return fail(SAPX_ERR_HIP, "pair template upload failed");
        }
    }
    return SAPX_OK;
}

static int32_t collide_params(const SapxPipeline* p, const float* body_q, const SapxContactsView* c,
                              CollideParams* P, PlaceParams* Q) {
    if (!p || !body_q || !c) return fail(SAPX_ERR_INVALID, "null argument");
    if (!c->count || !c->world_start || !c->shape0 || !c->shape1 || !c->point0 || !c->point1 || !c->normal ||
        !c->margin0 || !c->margin1)
        return fail(SAPX_ERR_INVALID, "contact view is missing arrays");
    if (c->capacity < 0) return fail(SAPX_ERR_INVALID, "negative contact capacity");
    const DevModel& dm = p->model->dm;
    std::memset(P, 0, sizeof(*P));
    std::memset(Q, 0, sizeof(*Q));
    P->m = dm;
    P->body_q = body_q;
    P->C = p->C;
    if (p->P >= 0) {
        P->m.t_pairs = p->t_pairs;
        P->m.P = p->P;
        P->m.has_convex = p->has_convex;
    }
    Q->W = dm.W;
    Q->C = p->C;
    Q->capacity = c->capacity;
    Q->world_start = c->world_start;
    Q->count = c->count;
    Q->shape0 = c->shape0; Q->shape1 = c->shape1;
    Q->p0 = c->point0; Q->p1 = c->point1; Q->nrm = c->normal;
    Q->m0 = c->margin0; Q->m1 = c->margin1;
    return SAPX_OK;
}

int32_t sapx_collide(SapxPipeline* p, const float* body_q, const SapxContactsView* c, void* stream) {
    CollideParams P;
    PlaceParams Q;
    const int32_t rc = collide_params(p, body_q, c, &P, &Q);
    if (rc != SAPX_OK) return rc;
    P.st_i = p->st_i;
    P.st_f = p->st_f;
    P.wcount = p->wcount;
    P.unsupported = p->unsupported;
    Q.st_i = p->st_i;
    Q.st_f = p->st_f;
    Q.wcount = p->wcount;
    HIP_TRY(launch_collide(P, Q, reinterpret_cast<hipStream_t>(stream)));
    return SAPX_OK;
}

int32_t sapx_collide_noalloc(SapxPipeline* p, const float* body_q, const SapxContactsView* c, void* stream) {
    CollideParams P;
    PlaceParams Q;
    const int32_t rc = collide_params(p, body_q, c, &P, &Q);
    if (rc != SAPX_OK) return rc;
    const size_t W = (size_t)p->model->dm.W;
    if (W > (size_t)kNoAllocWorlds || W * (size_t)p->C > kNoAllocSlots)
        return fail(SAPX_ERR_CAPACITY, "sapx_collide_noalloc takes at most " + std::to_string(kNoAllocWorlds) +
                                           " worlds and " + std::to_string(kNoAllocSlots) +
                                           " contact slots (worlds x max_contacts_per_world) per call");
    int* base = nullptr;
    HIP_TRY(collide_static_scratch(&base));
    P.unsupported = base;
    P.wcount = base + 64;
    P.st_i = base + 64 + kNoAllocWorlds;
    P.st_f = reinterpret_cast<float*>(base + 64 + kNoAllocWorlds + 2 * kNoAllocSlots);
    Q.wcount = P.wcount;
    Q.st_i = P.st_i;
    Q.st_f = P.st_f;
    HIP_TRY(launch_collide(P, Q, reinterpret_cast<hipStream_t>(stream)));
    return SAPX_OK;
}

static const WaveLayout& wave_layout(const SapxSolver* s, int cb) {
    return cb == 16 ? s->wl16 : (cb == 32 ? s->wl32 : s->wl64);
}

static int32_t batch_upload(SapxBatch* b, int slot) {
    const size_t G = b->s.size();
    SapxBatch::Views& v = b->views[slot];
    for (size_t g = 0; g < G; ++g) {
        SapxSolver* s = b->s[g];
        StepParams P;
        const int32_t rc = fill_step_params(s, &v.in[g], &v.out[g], &v.ctl[g], &v.c[g], v.dt, nullptr, nullptr, &P);
        if (rc != SAPX_OK) return rc;
        P.world_order = nullptr;
        P.learn_delta = nullptr;
        P.learn_target = nullptr;
        StepParams Pb = P;
        P.wl = wave_layout(s, b->small);
        P.wave_min_nc = 0;
        Pb.wl = s->wl64;
        Pb.wave_min_nc = b->small + 1;
        b->hparams[((size_t)slot * 2 + 0) * G + g] = P;
        b->hparams[((size_t)slot * 2 + 1) * G + g] = Pb;
        v.use_guess[g] = s->use_guess;
    }
    HIP_TRY(hipMemcpy(b->dparams + (size_t)slot * 2 * G, b->hparams.data() + (size_t)slot * 2 * G,
                      2 * G * sizeof(StepParams), hipMemcpyHostToDevice));
    v.bound = 1;
    return SAPX_OK;
}

int32_t sapx_batch_create(SapxSolver* const* solvers, int32_t count, SapxBatch** out) {
    if (!solvers || !out) return fail(SAPX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (count < 1 || count > kMaxBatchGroups)
        return fail(SAPX_ERR_CAPACITY, "a batch holds 1 to " + std::to_string(kMaxBatchGroups) + " solvers");
    for (int g = 0; g < count; ++g) {
        const SapxSolver* s = solvers[g];
        if (!s) return fail(SAPX_ERR_INVALID, "null solver");
        if (!s->wave_ok || s->npad > 48)
            return fail(SAPX_ERR_UNSUPPORTED, "batch steps need the one-wave kernel layout (at most 48 dofs per world)");
        if (s->learn_delta || s->learn_target)
            return fail(SAPX_ERR_UNSUPPORTED, "the learned warm start is not part of batch steps");
        if (!s->wl16.bytes) return fail(SAPX_ERR_UNSUPPORTED, "a solver's 16-slot layout does not fit the LDS");
    }
    SapxBatch* b = new SapxBatch();
    b->s.assign(solvers, solvers + count);
    b->hparams.resize((size_t)4 * count);
    for (auto& v : b->views) v.use_guess.assign(count, 0);
    if (hipMalloc(&b->dparams, (size_t)4 * count * sizeof(StepParams)) != hipSuccess ||
        hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return fail(SAPX_ERR_HIP, "batch allocation failed");
    }
    *out = b;
    return SAPX_OK;
}

void sapx_batch_destroy(SapxBatch* b) { delete b; }

int32_t sapx_batch_set_schedule(SapxBatch* b, int32_t small_capacity, int32_t fork) {
    if (!b) return fail(SAPX_ERR_INVALID, "null batch");
    if (small_capacity != 16 && small_capacity != 32 && small_capacity != 64)
        return fail(SAPX_ERR_INVALID, "small_capacity must be 16, 32 or 64");
    for (SapxSolver* s : b->s)
        if (!wave_layout(s, small_capacity).bytes || (small_capacity < s->opts.max_rigid_contact && !s->wl64.bytes))
            return fail(SAPX_ERR_UNSUPPORTED, "a solver's layout for this schedule does not fit the LDS");
    b->small = small_capacity;
    b->fork = fork ? 1 : 0;
    for (int slot = 0; slot < 2; ++slot)
        if (b->views[slot].bound) {
            const int32_t rc = batch_upload(b, slot);
            if (rc != SAPX_OK) return rc;
        }
    return SAPX_OK;
}

int32_t sapx_batch_bind(SapxBatch* b, int32_t slot, const SapxStateView* in, const SapxStateView* out,
                        const SapxControlView* ctl, const SapxContactsView* contacts, double dt) {
    if (!b || !in || !out || !ctl || !contacts) return fail(SAPX_ERR_INVALID, "null argument");
    if (slot < 0 || slot > 1) return fail(SAPX_ERR_INVALID, "slot must be 0 or 1");
    if (!(dt > 0.0)) return fail(SAPX_ERR_INVALID, "dt must be positive");
    const size_t G = b->s.size();
    for (size_t g = 0; g < G; ++g)
        if (!contacts[g].world_start || !contacts[g].count)
            return fail(SAPX_ERR_INVALID, "batch steps take world-grouped contacts (world_start)");
    SapxBatch::Views& v = b->views[slot];
    v.in.assign(in, in + G);
    v.out.assign(out, out + G);
    v.ctl.assign(ctl, ctl + G);
    v.c.assign(contacts, contacts + G);
    v.dt = dt;
    return batch_upload(b, slot);
}

int32_t sapx_batch_step(SapxBatch* b, int32_t slot, void* stream) {
    if (!b) return fail(SAPX_ERR_INVALID, "null batch");
    if (slot < 0 || slot > 1 || !b->views[slot].bound) return fail(SAPX_ERR_INVALID, "slot is not bound");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const size_t G = b->s.size();
    // a solver whose warm start became valid since the upload (its first step): refresh the slot
    bool stale = false;
    for (size_t g = 0; g < G; ++g) stale |= b->views[slot].use_guess[g] != b->s[g]->use_guess;
    if (stale) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(SAPX_ERR_INVALID, "step the batch once before capturing it (warm-start state changes)");
        HIP_TRY(hipStreamSynchronize(st));
        const int32_t rc = batch_upload(b, slot);
        if (rc != SAPX_OK) return rc;
    }
    BatchMap M0{}, M1{};
    int blocks = 0;
    size_t smem0 = 0, smem1 = 0;
    bool large = false;
    for (size_t g = 0; g < G; ++g) {
        const SapxSolver* s = b->s[g];
        M0.block_start[g] = blocks;
        M0.npad[g] = s->npad;
        blocks += s->model->dm.W;
        smem0 = std::max(smem0, (size_t)wave_layout(s, b->small).bytes);
        smem1 = std::max(smem1, (size_t)s->wl64.bytes);
        large |= b->small < 64 && s->opts.max_rigid_contact > b->small;
    }
    M0.block_start[G] = blocks;
    M0.groups = (int)G;
    M1 = M0;
    M0.params = b->dparams + (size_t)slot * 2 * G;
    M1.params = b->dparams + ((size_t)slot * 2 + 1) * G;
    auto launch = [&](const BatchMap& M, int cb, size_t smem, hipStream_t sq) {
        return cb == 16 ? launch_batch_16(M, blocks, smem, sq)
                        : (cb == 32 ? launch_batch_32(M, blocks, smem, sq) : launch_batch_64(M, blocks, smem, sq));
    };
    if (!large) {
        HIP_TRY(launch(M0, b->small, smem0, st));
    } else if (b->fork) {
        HIP_TRY(hipEventRecord(b->ev_fork, st));
        HIP_TRY(hipStreamWaitEvent(b->aux, b->ev_fork, 0));
        HIP_TRY(launch(M1, 64, smem1, b->aux));
        HIP_TRY(hipEventRecord(b->ev_join, b->aux));
        HIP_TRY(launch(M0, b->small, smem0, st));
        HIP_TRY(hipStreamWaitEvent(st, b->ev_join, 0));
    } else {
        HIP_TRY(launch(M0, b->small, smem0, st));
        HIP_TRY(launch(M1, 64, smem1, st));
    }
    for (SapxSolver* s : b->s) s->use_guess = 1;
    return SAPX_OK;
}

}  // extern "C"
