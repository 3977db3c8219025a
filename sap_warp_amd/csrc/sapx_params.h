// Kernel parameter blocks shared by the host API and the kernels.
#pragma once

#include "sapx_device.h"

struct SapxModel;

namespace sapx {

// LDS layout of the one-wave-per-world step kernel (byte offsets); js / ls are
// the odd row strides of the contact Jacobian and of the factor rows.
struct WaveLayout {
    int o_topo, o_mp, o_q, o_bq, o_bqd, o_S, o_vec, o_ck, o_cG, o_A, o_J, o_Lt;
    int bytes;
    int js, ls;
};

struct StepParams {
    DevModel m;
    const float* jq_in;
    const float* jqd_in;
    const float* bq_in;
    const float* bf_in;
    const float* c_jf;
    const float* c_tpos;
    const float* c_tvel;
    const float* c_act;
    const int* cworld_start;
    const int* cshape0;
    const int* cshape1;
    const float* cpoint0;
    const float* cpoint1;
    const float* cnormal;
    const float* cmargin0;
    const float* cmargin1;
    float* jq_out;
    float* jqd_out;
    float* bq_out;
    float* bqd_out;
    float* bf_out;
    double* v_guess;
    int use_guess;
    int* st_iters;
    int* st_ls;
    int* st_nc;
    int* st_flags;
    unsigned long long* st_hist;  // 256 iteration bins + [256] line-search total + [257] world-steps
    double* dbg_v;
    double* dbg_vstar;
    double* dbg_v0;
    double* dbg_gamma;
    double* dbg_phi0;
    double* dbg_w;
    double* dbg_adiag;
    double* dbg_A;      // W x n x n: dynamics matrix A = H + diag(armature) as the solve uses it
    double* dbg_J;      // W x C x 3 x n: contact Jacobian rows (t1, t2, n per slot; zero past the count)
    double dt;
    const double* dt_world;
    int C;
    double fallback_mu, fallback_k, fallback_tau, beta, sigma, diag_shift;
    int max_iter;
    double opt_abs, opt_rel, cost_abs, cost_rel;
    int ls_mode, ls_max_iter;
    double armijo_c, rho, rel_slop;
    int weight_mode, point_mode, integ_mode;
    int npad, js, ast;  // padded dof count, J row stride, A row stride (both odd)
    // LDS byte offsets
    int o_A, o_H, o_J, o_bq, o_S, o_ck, o_cv, o_cG, o_cb, o_dv64, o_dv32, o_dflag, o_red, o_q, o_misc, o_bqd, o_topo;
    int smem_bytes;
    // large worlds (fused kernel): A, H (with the contact set-up temporaries) and J (with the
    // kinematics scratch) live in a per-world block of device memory instead of LDS; g_work is
    // null when they fit in LDS.  Offsets and stride in floats.
    float* g_work;
    size_t g_stride;
    size_t g_A, g_H, g_J;
    WaveLayout wl;       // one-wave kernel layout for this launch's contact variant
    int wave_min_nc;     // the variant processes worlds with wave_min_nc <= nc <= its capacity
    const int* world_order;  // blockIdx -> world (a permutation of 0..W-1), or null for the identity
    // learned warm start (wave kernels): predicted correction of the plain start (W x N, free-joint
    // rows in the joint's body frame) and the training target written back (same layout)
    const float* learn_delta;
    float* learn_target;
};

// input features of the learned warm start (sapx_learned.hip)
struct FeatureParams {
    DevModel m;
    const float* jq;
    const float* jqd;
    const float* bq;
    const float* tpos;  // may be null (zero targets)
    const float* tvel;
    const int* cworld_start;  // may be null (no contacts)
    const int* cshape0;
    const int* cshape1;
    const float* cpoint0;
    const float* cpoint1;
    const float* cnormal;
    const float* cmargin0;
    const float* cmargin1;
    int C;                // contact capacity per world
    const int* feat_off;  // J+1: first feature of each joint's block (bodies follow at feat_off[J])
    int frame_qs;         // local coordinate start of the first free joint, -1 if none
    const float* prev_target;  // W x N previous correction, or null (zeros)
    int F;                // features per world
    float* out;           // W x F
};

// Collision: one workgroup per world stages its contacts in LDS, then copies them (coalesced) into
// the world's block of a W x C staging area with its count; the place kernel moves every world's
// block to its offset in the flat, world-grouped contact arrays.
// Batch step: the worlds of up to kMaxBatchGroups models (kinds of different dof counts) in one
// launch; block b belongs to the group g with block_start[g] <= b < block_start[g + 1].
constexpr int kMaxBatchGroups = 4;
struct BatchMap {
    const StepParams* params;  // device, one per group
    int groups;
    int block_start[kMaxBatchGroups + 1];
    int npad[kMaxBatchGroups];
};

// sapx_collide_noalloc: limits of its static scratch (55 MB per device)
constexpr int kNoAllocWorlds = 65536;          // worlds per call
constexpr size_t kNoAllocSlots = 1u << 20;     // worlds x contacts per world per call
constexpr size_t kNoAllocWords = 64 + kNoAllocWorlds + 13 * kNoAllocSlots;
constexpr int kNoAllocMaxDevices = 64;

constexpr int kCollideThreads = 128;
constexpr size_t kCollideMaxLds = 64 * 1024;  // dynamic LDS of one collide workgroup

struct CollideParams {
    DevModel m;
    const float* body_q;
    int C;          // contact capacity per world
    int* st_i;      // W x (2 C): shape0[C], shape1[C] per world
    float* st_f;    // W x (11 C): margin0[C], margin1[C], point0[3C], point1[3C], normal[3C] per world
    int* wcount;    // W: staged contacts per world
    int* unsupported;
};

struct PlaceParams {
    int W, C;
    int capacity;  // slots of the flat contact arrays
    const int* wcount;
    const int* st_i;
    const float* st_f;
    int* world_start;  // W + 1
    int* count;
    int* shape0;
    int* shape1;
    float* p0;
    float* p1;
    float* nrm;
    float* m0;
    float* m1;
};

// grouping of an arbitrary flat contact list by world (stable radix sort)
// One large world's collision (thread per candidate pair): shape tables, the pair list, scratch
// and the flat contact output (sapx_bigworld.hip drives it)
constexpr int kBigSlots = 5;  // contact slots staged per candidate pair
struct BigCollideParams {
    int nshapes, P, C;
    const int* shape_type;
    const int* shape_body;
    const float* shape_transform;  // x7
    const float* shape_scale;      // x3
    const float* shape_radius;
    const float* shape_margin;
    const float* shape_gap;
    const int* pairs;  // P x 2 shape indices, candidate order
    const float* body_q;
    float* X;        // nshapes x 8 world transforms
    float* AB;       // nshapes x 6 AABBs
    int* cnt;        // P: contacts passing the gap test
    int* cvx;        // P: 1 for a convex (per-contact capacity) pair
    int* off;        // P: exclusive prefix of cnt (written by the caller's scan)
    int* st_sh;      // P x kBigSlots x 2 shape ids
    float* st_f;     // P x kBigSlots x 11: point0, point1, normal, margin0, margin1
    int* unsupported;
    int* out_shape0;
    int* out_shape1;
    float* out_p0;
    float* out_p1;
    float* out_n;
    float* out_m0;
    float* out_m1;
    int* out_pair;     // candidate index of each contact (may be null)
    int* world_start;  // 2
    int* count;        // 1
};
const DevModel& model_dev(const ::SapxModel* m);
hipError_t launch_big_pairs(const BigCollideParams& B, hipStream_t stream);
hipError_t launch_big_write(const BigCollideParams& B, hipStream_t stream);

struct GroupParams {
    int W, NB, capacity;
    const int* count;
    const int* shape_body;
    const int* in_shape0;
    const int* in_shape1;
    const float* in_p0;
    const float* in_p1;
    const float* in_n;
    const float* in_m0;
    const float* in_m1;
    int* keys;
    int* vals;
    int* keys_out;
    int* vals_out;
    int* world_start;
    int* out_shape0;
    int* out_shape1;
    float* out_p0;
    float* out_p1;
    float* out_n;
    float* out_m0;
    float* out_m1;
    void* cub_tmp;
    size_t cub_tmp_bytes;
};

}  // namespace sapx
