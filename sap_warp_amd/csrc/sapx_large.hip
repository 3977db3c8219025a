// SAP Newton solve of one large world on the whole GPU (gfx950).
//
// The step kernels give each world one wave or one workgroup; a single world of thousands
// of dofs and tens of thousands of contacts (a bin of free bodies around an articulated
// robot) needs every stage spread over the GPU instead.  This solver takes the problem a
// world's free motion and contact Jacobian produced -- the dynamics matrix A, v*, the
// contact rows in a sparse per-contact layout (each contact's three rows over at most kmax
// dofs: the dofs of the two bodies' articulations it touches), the contact parameters and
// the PD / limit terms -- and runs the reference's Newton iteration (contact_solve.py:
// 6129-6285) with its projection (743-921), optimality and cost tests (1970-2087), exact
// root (2913-3140) and Armijo (2670-2905) line searches:
//
//   evaluate      one thread per contact: J v, projection, gamma, G, cost, J^T gamma
//                 (float64 atomics); one wave per dof row: A r, A v, the gradient
//   Hessian       A + diag(model terms) copied, then one workgroup per contact adds
//                 J_c^T G_c J_c into the dense float32 Newton matrix
//   direction     the blocked Cholesky of sapx_chol.hip (panel / MFMA Schur kernels over the
//                 whole GPU) and its float64-refined solve
//   line search   J dv per contact and A dv per row, then one 1024-thread workgroup runs
//                 every trial of the search on the device (contacts and dofs in strides,
//                 block reductions of cost and its two derivatives)
//
// The host drives the iterations and reads back one small record per iteration (cost,
// norms and the previous line search's outcome), so a solve is not graph-capturable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/sapx.h"
#include "sapx_kcommon.h"

namespace sapx {
void set_last_error(const char* msg);
}

namespace {

using sapx::clamp_a;
using sapx::clamp_d;
using sapx::clamp_s;
using sapx::project_contact;

constexpr int NT = 256;
constexpr int NR = 1024;                 // threads of the reduction / line-search workgroup
constexpr double kPi = 3.14159265358979323846;
constexpr double kExactAlphaMax = 1.5;   // contact_solve.py:52
constexpr double kExactFTol = 1.0e-8;    // contact_solve.py:54
constexpr int kRedSlots = 8;

int32_t lfail(int32_t code, const std::string& msg) {
    sapx::set_last_error(msg.c_str());
    return code;
}

struct ModelTerms {
    const int* pd_active;
    const double* pd_a;
    const double* pd_gain;
    const double* pd_limit;
    const int* lo_active;
    const int* up_active;
    const double* lo_vhat;
    const double* up_vhat;
    const double* lo_r;
    const double* up_r;
    double dt;
};

// PD + limit terms of dof i at velocity vi (contact_solve.py:1569-1612): returns the cost,
// writes the impulse and the Hessian diagonal
__device__ __forceinline__ double model_terms(const ModelTerms& m, int i, double vi, double* imp, double* hd) {
    double im = 0.0, h = 0.0, cost = 0.0;
    if (m.pd_active) {
        if (m.pd_active[i]) {
            const double gain = m.pd_gain[i], lim = m.pd_limit[i];
            const double y = m.pd_a[i] - gain * vi;
            im += m.dt * clamp_s(y, lim);
            h += m.dt * gain * clamp_d(y, lim);
            if (gain > 0.0) cost += (m.dt / gain) * clamp_a(y, lim);
        }
        if (m.lo_active[i]) {
            const double g = (m.lo_vhat[i] - vi) / m.lo_r[i];
            if (g > 0.0) {
                im += g;
                h += 1.0 / m.lo_r[i];
                cost += 0.5 * m.lo_r[i] * g * g;
            }
        }
        if (m.up_active[i]) {
            const double g = (m.up_vhat[i] + vi) / m.up_r[i];
            if (g > 0.0) {
                im -= g;
                h += 1.0 / m.up_r[i];
                cost += 0.5 * m.up_r[i] * g * g;
            }
        }
    }
    *imp = im;
    *hd = h;
    return cost;
}

// three rows of contact c against x: out[r] = sum_k J[r][k] x[col_k] (entries outside 0..n-1
// are unused slots)
__device__ __forceinline__ void jrows(const int* cols, const float* vals, int kmax, int n, const double* x,
                                      double out[3]) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < kmax; ++k) {
        const int col = cols[k];
        if (col < 0 || col >= n) continue;
        const double xv = x[col];
        s0 += (double)vals[k] * xv;
        s1 += (double)vals[kmax + k] * xv;
        s2 += (double)vals[2 * kmax + k] * xv;
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
}

// block-wide sums of three doubles (NR threads); every thread gets the totals
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* sm) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) {
        sm[3 * w] = a;
        sm[3 * w + 1] = b;
        sm[3 * w + 2] = c;
    }
    __syncthreads();
    a = b = c = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        a += sm[3 * k];
        b += sm[3 * k + 1];
        c += sm[3 * k + 2];
    }
}

// contact constants (contact_solve.py:791-822): rt, rn, vhat, mu, 1/rt, 1/rn, mu rt/rn,
// 1 / (1 + mu^2 rt / rn)
__global__ __launch_bounds__(NT) void l_constants(int nc, const double* phi0, const double* weight,
                                                  const double* stiffness, const double* tau, const double* mu_in,
                                                  double dt, double beta, double sigma, double* kc) {
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= nc) return;
    double w = weight[c];
    if (w < 1.0e-12 || !isfinite(w)) w = 1.0e-12;
    double k = stiffness[c];
    if (k <= 0.0 || !isfinite(k)) k = 1.0;
    const double tau_c = fmax(tau[c], 0.0);
    const double bf = beta * beta / (4.0 * kPi * kPi);
    const double rn = fmax(fmax(bf * w, 1.0 / (dt * k * (dt + tau_c))), 1.0e-30);
    const double rt = fmax(sigma * w, 1.0e-30);
    double mu = mu_in[c];
    if (mu < 0.0 || !isfinite(mu)) mu = 0.0;
    double* k8 = kc + (size_t)c * 8;
    k8[0] = rt;
    k8[1] = rn;
    k8[2] = -phi0[c] / (dt + tau_c);
    k8[3] = mu;
    k8[4] = 1.0 / rt;
    k8[5] = 1.0 / rn;
    const double mu_tilde = mu * sqrt(rt / rn);
    k8[6] = mu * rt / rn;
    k8[7] = 1.0 / (1.0 + mu_tilde * mu_tilde);
}

// start: v = guess on participating dofs, v* elsewhere; D = diag(A)^-1/2 (clamped)
__global__ __launch_bounds__(NT) void l_init(int n, const float* A, const double* v_star, const double* v_guess,
                                             const int* part, double* v, double* dsc) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    v[i] = part[i] ? v_guess[i] : v_star[i];
    double d = (double)A[(size_t)i * n + i];
    if (!(d >= 1.0e-12) || !isfinite(d)) d = 1.0e-12;
    dsc[i] = 1.0 / sqrt(d);
}

// model terms per dof; jimp starts at the model impulse (contacts add J^T gamma)
__global__ __launch_bounds__(NT) void l_model(int n, ModelTerms m, const double* v, double* jimp, double* hd,
                                              double* mcost) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    double imp, h;
    mcost[i] = model_terms(m, i, v[i], &imp, &h);
    jimp[i] = imp;
    hd[i] = h;
}

// per contact: J v, projection, gamma, G, cost; J^T gamma added to jimp
__global__ __launch_bounds__(NT) void l_contacts(int nc, int kmax, int n, const int* Jc, const float* Jv,
                                                 const double* kc, const double* v, double* gam, float* G, double* ccost,
                                                 double* jimp) {
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= nc) return;
    const int* cols = Jc + (size_t)c * kmax;
    const float* vals = Jv + (size_t)c * 3 * kmax;
    double vc[3], g[3];
    jrows(cols, vals, kmax, n, v, vc);
    ccost[c] = project_contact(kc + (size_t)c * 8, vc, g, G + (size_t)c * 9, true);
    gam[3 * c] = g[0];
    gam[3 * c + 1] = g[1];
    gam[3 * c + 2] = g[2];
    for (int k = 0; k < kmax; ++k) {
        const int col = cols[k];
        if (col < 0 || col >= n) continue;
        const double t = (double)vals[k] * g[0] + (double)vals[kmax + k] * g[1] + (double)vals[2 * kmax + k] * g[2];
        if (t != 0.0) atomicAdd(jimp + col, t);
    }
}

// one wave per row: A (v - v*), A v, gradient; r_i (A r)_i for the quadratic cost
__global__ __launch_bounds__(NT) void l_rows(int n, const float* A, const double* v, const double* v_star,
                                             const double* jimp, double* Av, double* grad, double* rAr) {
    const int i = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const float* a = A + (size_t)i * n;
    double s_av = 0.0, s_as = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double aij = (double)a[j];
        s_av += aij * v[j];
        s_as += aij * v_star[j];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s_av += __shfl_xor(s_av, off, 64);
        s_as += __shfl_xor(s_as, off, 64);
    }
    if (lane == 0) {
        const double ar = s_av - s_as;
        Av[i] = s_av;
        grad[i] = ar - jimp[i];
        rAr[i] = (v[i] - v_star[i]) * ar;
    }
}

// evaluation record: cost, ||D grad||, tolerance (contact_solve.py:1970-2087)
__global__ __launch_bounds__(NR) void l_eval_reduce(int n, int nc, const double* rAr, const double* ccost,
                                                    const double* mcost, const double* grad, const double* Av,
                                                    const double* jimp, const double* dsc, const int* part,
                                                    double abs_tol, double rel_tol, double* red) {
    __shared__ double sm[3 * (NR / 64)];
    double cost = 0.0, g2 = 0.0, p2 = 0.0, j2 = 0.0, dummy = 0.0, dummy2 = 0.0;
    for (int i = threadIdx.x; i < n; i += NR) {
        cost += 0.5 * rAr[i] + mcost[i];
        if (part[i]) {
            const double d = dsc[i];
            g2 += (d * grad[i]) * (d * grad[i]);
            p2 += (d * Av[i]) * (d * Av[i]);
            j2 += (d * jimp[i]) * (d * jimp[i]);
        }
    }
    for (int c = threadIdx.x; c < nc; c += NR) cost += ccost[c];
    block_sum3(cost, g2, p2, sm);
    block_sum3(j2, dummy, dummy2, sm);
    if (threadIdx.x == 0) {
        red[0] = cost;
        red[1] = sqrt(g2);
        red[2] = abs_tol + rel_tol * fmax(sqrt(p2), sqrt(j2));
    }
}

// H = A + diag(hd + shift) (float32)
__global__ __launch_bounds__(NT) void l_hinit(int n, const float* A, const double* hd, double shift, float* H) {
    const size_t total = (size_t)n * n;
    for (size_t e = (size_t)blockIdx.x * NT + threadIdx.x; e < total; e += (size_t)gridDim.x * NT) {
        const int i = (int)(e / n), j = (int)(e - (size_t)i * n);
        float h = A[e];
        if (i == j) h = (float)((double)h + hd[i] + shift);
        H[e] = h;
    }
}

// one 64-thread workgroup per contact: H += J_c^T G_c J_c over the contact's dofs
__global__ __launch_bounds__(64) void l_hcontacts(int n, int kmax, const int* Jc, const float* Jv, const float* G,
                                                  float* H) {
    extern __shared__ float ls[];  // 3 kmax: G J; kmax: cols (as float bits)
    const int c = blockIdx.x;
    const int* cols = Jc + (size_t)c * kmax;
    const float* vals = Jv + (size_t)c * 3 * kmax;
    const float* Gc = G + (size_t)c * 9;
    float* GJ = ls;
    int* cl = reinterpret_cast<int*>(ls + 3 * kmax);
    for (int k = threadIdx.x; k < kmax; k += 64) {
        cl[k] = cols[k];
        const float j0 = vals[k], j1 = vals[kmax + k], j2 = vals[2 * kmax + k];
#pragma unroll
        for (int r = 0; r < 3; ++r) GJ[r * kmax + k] = Gc[3 * r] * j0 + Gc[3 * r + 1] * j1 + Gc[3 * r + 2] * j2;
    }
    __syncthreads();
    const int kk = kmax * kmax;
    for (int e = threadIdx.x; e < kk; e += 64) {
        const int k1 = e / kmax, k2 = e - k1 * kmax;
        const int c1 = cl[k1], c2 = cl[k2];
        if (c1 < 0 || c2 < 0 || c1 >= n || c2 >= n) continue;
        const float val = vals[k1] * GJ[k2] + vals[kmax + k1] * GJ[kmax + k2] + vals[2 * kmax + k1] * GJ[2 * kmax + k2];
        if (val != 0.0f) atomicAdd(H + (size_t)c1 * n + c2, val);
    }
}

__global__ __launch_bounds__(NT) void l_neg(int n, const double* x, double* y) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n) y[i] = -x[i];
}

// line-search data: J dv per contact
__global__ __launch_bounds__(NT) void l_dvc(int nc, int kmax, int n, const int* Jc, const float* Jv,
                                            const double* dv, const double* v, double* dvc, double* vc) {
    const int c = blockIdx.x * NT + threadIdx.x;
    if (c >= nc) return;
    double d[3], x[3];
    jrows(Jc + (size_t)c * kmax, Jv + (size_t)c * 3 * kmax, kmax, n, dv, d);
    jrows(Jc + (size_t)c * kmax, Jv + (size_t)c * 3 * kmax, kmax, n, v, x);
    dvc[3 * c] = d[0];
    dvc[3 * c + 1] = d[1];
    dvc[3 * c + 2] = d[2];
    vc[3 * c] = x[0];
    vc[3 * c + 1] = x[1];
    vc[3 * c + 2] = x[2];
}

// line-search data per row: p = A dv; partials grad.dv, p.(v - v*), dv.p
__global__ __launch_bounds__(NT) void l_rows_dv(int n, const float* A, const double* dv, const double* v,
                                                const double* v_star, const double* grad, double* part3) {
    const int i = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const float* a = A + (size_t)i * n;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += (double)a[j] * dv[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        part3[3 * i] = grad[i] * dv[i];
        part3[3 * i + 1] = s * (v[i] - v_star[i]);
        part3[3 * i + 2] = dv[i] * s;
    }
}

struct LsArgs {
    int n, nc, variant, max_it;
    double armijo_c, rho, relative_slop, cost_abs_tol, cost_rel_tol;
    const double *kc, *vc, *dvc, *v, *dv, *part3, *red;
    ModelTerms m;
    double* out;  // accepted, alpha, trials, status
};

// cost and its derivatives along dv at alpha (contact_solve.py:2234-2420)
__device__ void ls_trial(const LsArgs& a, double alpha, bool second, double m0, double da0, double d2a, double* sm,
                         double* cost, double* d1, double* d2) {
    double cs = 0.0, g1 = 0.0, g2 = 0.0;
    for (int c = threadIdx.x; c < a.nc; c += NR) {
        double vc[3], dvc[3], g[3], cc;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            dvc[r] = a.dvc[3 * c + r];
            vc[r] = a.vc[3 * c + r] + alpha * dvc[r];
        }
        double curv = 0.0;
        if (second) {
            curv = sapx::project_curvature(a.kc + (size_t)c * 8, vc, dvc, g, &cc);
        } else {
            float Gd[9];
            cc = project_contact(a.kc + (size_t)c * 8, vc, g, Gd, false);
        }
        cs += cc;
        g1 -= dvc[0] * g[0] + dvc[1] * g[1] + dvc[2] * g[2];
        g2 += curv;
    }
    for (int i = threadIdx.x; i < a.n; i += NR) {
        const double dvi = a.dv[i];
        double imp, hd;
        cs += model_terms(a.m, i, a.v[i] + alpha * dvi, &imp, &hd);
        g1 -= dvi * imp;
        g2 += hd * dvi * dvi;
    }
    block_sum3(cs, g1, g2, sm);
    *cost = m0 + alpha * da0 + 0.5 * alpha * alpha * d2a + cs;
    *d1 = da0 + alpha * d2a + g1;
    *d2 = d2a + g2;
}

// the whole line search in one workgroup: exact root (contact_solve.py:2913-3140) or
// Armijo backtracking (2670-2905); writes (accepted, alpha, trials, status)
__global__ __launch_bounds__(NR) void l_line_search(LsArgs a) {
    __shared__ double sm[3 * (NR / 64)];
    double dell0 = 0.0, da0 = 0.0, d2a = 0.0;
    for (int i = threadIdx.x; i < a.n; i += NR) {
        dell0 += a.part3[3 * i];
        da0 += a.part3[3 * i + 1];
        d2a += a.part3[3 * i + 2];
    }
    block_sum3(dell0, da0, d2a, sm);
    const double ell0 = a.red[0];
    // m0 = 1/2 r^T A r: the quadratic part of the cost at alpha = 0
    const double m0 = a.red[3];
    const bool t0 = threadIdx.x == 0;
    double accepted = 0.0, alpha = 1.0, trials = 0.0, status = 0.0;
    if (a.variant == SAPX_LS_EXACT_ROOT) {
        const double amax = kExactAlphaMax, tol = kExactFTol;
        bool done = false;
        if (!(dell0 < 0.0) || !isfinite(dell0)) {
            alpha = amax;
            status = -1.0;
            done = true;
        }
        double ell, d1, d2;
        if (!done) {
            ls_trial(a, amax, true, m0, da0, d2a, sm, &ell, &d1, &d2);
            if (d1 <= 0.0) {
                accepted = 1.0;
                alpha = amax;
                done = true;
            } else if (-dell0 < a.cost_abs_tol + a.cost_rel_tol * ell) {
                accepted = 1.0;
                alpha = 1.0;
                done = true;
            }
        }
        double scale = -dell0, f_hi = 0.0, root = 0.0, x_lo = 0.0, x_hi = amax, minus_dx = -amax,
               minus_dx_prev = -amax, x_tol = 0.0;
        if (!done) {
            f_hi = d1 / scale;
            if (!isfinite(f_hi) || f_hi <= 0.0) {
                alpha = amax;
                status = -3.0;
                done = true;
            }
        }
        if (!done) {
            double guess = -dell0 / d2;
            const double h = amax;
            const double theta = dell0 + d1 - 3.0 * (ell - ell0) / h;
            const double disc = theta * theta - dell0 * d1;
            if (isfinite(disc) && disc >= 0.0) {
                const double gm = sqrt(disc);
                const double denom = d1 - dell0 + 2.0 * gm;
                if (denom != 0.0) {
                    const double cubic = h - h * (d1 + gm - theta) / denom;
                    if (isfinite(cubic) && 0.0 < cubic && cubic <= h) guess = cubic;
                }
            }
            if (guess > amax) guess = amax;
            if (!isfinite(guess) || guess < 0.0 || guess > amax || tol * guess <= 0.0) {
                alpha = amax;
                status = -4.0;
                done = true;
            }
            root = guess;
            x_tol = tol * guess;
        }
        int iteration = 1;
        while (!done) {
            double cst, dd1, dd2;
            ls_trial(a, root, true, m0, da0, d2a, sm, &cst, &dd1, &dd2);
            const double f = dd1 / scale, df = dd2 / scale;
            if ((f < 0.0 && f_hi >= 0.0) || (f >= 0.0 && f_hi < 0.0)) {
                x_lo = root;
            } else {
                x_hi = root;
                f_hi = f;
            }
            if (fabs(f) < tol) {
                accepted = 1.0;
                alpha = root;
                trials = iteration;
                break;
            }
            if (iteration >= a.max_it) {
                alpha = root;
                trials = iteration;
                status = -5.0;
                break;
            }
            const bool slow = 2.0 * fabs(f) > fabs(minus_dx_prev * df);
            minus_dx_prev = minus_dx;
            const double bis_dx = 0.5 * (x_lo - x_hi);
            const double bis_root = x_lo - bis_dx;
            double newton_dx = 0.0, newton_root = root;
            if (fabs(df) > 0.0) {
                newton_dx = f / df;
                newton_root = root - newton_dx;
            }
            const bool use_bisect = slow || newton_root < x_lo || newton_root > x_hi || !isfinite(newton_root) ||
                                    !isfinite(newton_dx);
            const double nxt = use_bisect ? bis_root : newton_root;
            const double ndx = use_bisect ? bis_dx : newton_dx;
            if (fabs(ndx) < x_tol) {
                accepted = 1.0;
                alpha = nxt;
                trials = iteration;
                break;
            }
            root = nxt;
            minus_dx = ndx;
            ++iteration;
        }
    } else {
        // Armijo backtracking
        const double amax = 1.0 / a.rho;
        bool done = false;
        if (!(dell0 < 0.0) || !isfinite(dell0)) {
            alpha = 1.0;
            status = -1.0;
            done = true;
        }
        double ell = 0.0, d1 = 0.0, d2 = 0.0, slop = 0.0, a_prev = 0.0, e_prev = 0.0;
        if (!done) {
            alpha = amax;
            ls_trial(a, alpha, false, m0, da0, d2a, sm, &ell, &d1, &d2);
            const double scale = fmax(1.0, 0.5 * (fabs(ell) + fabs(ell0)));
            slop = (a.relative_slop / 10.0) * scale;
            if (d1 < 0.0 || d1 < slop) {
                accepted = 1.0;
                done = true;
            } else {
                a_prev = alpha;
                e_prev = ell;
                if (a.max_it <= 1) {
                    status = -2.0;
                    done = true;
                }
            }
        }
        int iteration = 1;
        while (!done) {
            alpha *= a.rho;
            ls_trial(a, alpha, false, m0, da0, d2a, sm, &ell, &d1, &d2);
            const double denom = alpha - a_prev;
            const double dell_approx = fabs(denom) > 0.0 ? (ell - e_prev) / denom : 0.0;
            const bool ok_a = ell < ell0 + a.armijo_c * alpha * dell0;
            if (fabs(dell_approx) < slop) {
                accepted = 1.0;
                trials = iteration;
                break;
            }
            if (ell > e_prev && ok_a) {
                const bool ok_prev = e_prev < ell0 + a.armijo_c * a_prev * dell0;
                if (ok_prev) alpha = a_prev;
                accepted = 1.0;
                trials = iteration;
                break;
            }
            if (iteration >= a.max_it - 1) {
                if (ok_a) {
                    accepted = 1.0;
                    trials = a.max_it;
                } else {
                    trials = iteration;
                    status = -2.0;
                }
                break;
            }
            a_prev = alpha;
            e_prev = ell;
            ++iteration;
        }
    }
    if (t0) {
        a.out[0] = accepted;
        a.out[1] = alpha;
        a.out[2] = trials;
        a.out[3] = status;
    }
}

// v += alpha dv when the search accepted
__global__ __launch_bounds__(NT) void l_update(int n, const double* ls, const double* dv, double* v) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n && ls[0] != 0.0) v[i] += ls[1] * dv[i];
}

// m0 = 1/2 r^T A r into red[3] (cost at alpha = 0 minus the regularisers)
__global__ __launch_bounds__(NR) void l_m0(int n, const double* rAr, double* red) {
    __shared__ double sm[3 * (NR / 64)];
    double s = 0.0, z1 = 0.0, z2 = 0.0;
    for (int i = threadIdx.x; i < n; i += NR) s += 0.5 * rAr[i];
    block_sum3(s, z1, z2, sm);
    if (threadIdx.x == 0) red[3] = s;
}

__global__ __launch_bounds__(NT) void l_copy(int n, const double* x, double* y) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n) y[i] = x[i];
}

}  // namespace

struct SapxLarge {
    int max_n = 0, max_nc = 0, kmax = 0;
    SapxChol* chol = nullptr;
    float* H = nullptr;
    double *v = nullptr, *dsc = nullptr, *grad = nullptr, *jimp = nullptr, *Av = nullptr, *rAr = nullptr, *hd = nullptr,
           *mcost = nullptr, *dv = nullptr, *rhs = nullptr, *part3 = nullptr;
    double *kc = nullptr, *gam = nullptr, *ccost = nullptr, *vc = nullptr, *dvc = nullptr;
    float* G = nullptr;
    double* red = nullptr;   // device: cost, gn, tol, m0, then the line search record (4)
    double* hred = nullptr;  // pinned host copy
};

extern "C" {

int32_t sapx_large_create(int32_t max_dofs, int32_t max_contacts, int32_t kmax, SapxLarge** out) {
    if (!out) return lfail(SAPX_ERR_INVALID, "sapx_large_create: null output");
    *out = nullptr;
    if (max_dofs <= 0 || max_contacts < 0 || kmax <= 0 || kmax > 256)
        return lfail(SAPX_ERR_INVALID, "sapx_large_create: max_dofs > 0, max_contacts >= 0, 1 <= kmax <= 256");
    SapxLarge* s = new SapxLarge();
    s->max_n = max_dofs;
    s->max_nc = max_contacts;
    s->kmax = kmax;
    const int32_t rc = sapx_chol_create(max_dofs, 1, &s->chol);
    if (rc != SAPX_OK) {
        delete s;
        return rc;
    }
    const size_t n = (size_t)max_dofs, nc = (size_t)(max_contacts > 0 ? max_contacts : 1);
    bool ok = hipMalloc(&s->H, n * n * sizeof(float)) == hipSuccess;
    for (double** p : {&s->v, &s->dsc, &s->grad, &s->jimp, &s->Av, &s->rAr, &s->hd, &s->mcost, &s->dv, &s->rhs})
        ok = ok && hipMalloc(p, n * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&s->part3, 3 * n * sizeof(double)) == hipSuccess &&
         hipMalloc(&s->kc, 8 * nc * sizeof(double)) == hipSuccess &&
         hipMalloc(&s->gam, 3 * nc * sizeof(double)) == hipSuccess &&
         hipMalloc(&s->ccost, nc * sizeof(double)) == hipSuccess &&
         hipMalloc(&s->vc, 3 * nc * sizeof(double)) == hipSuccess &&
         hipMalloc(&s->dvc, 3 * nc * sizeof(double)) == hipSuccess &&
         hipMalloc(&s->G, 9 * nc * sizeof(float)) == hipSuccess &&
         hipMalloc(&s->red, kRedSlots * sizeof(double)) == hipSuccess &&
         hipHostMalloc(&s->hred, kRedSlots * sizeof(double)) == hipSuccess &&
         hipMemset(s->red, 0, kRedSlots * sizeof(double)) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        sapx_large_destroy(s);
        return lfail(SAPX_ERR_HIP, "sapx_large_create: device allocation failed");
    }
    *out = s;
    return SAPX_OK;
}

void sapx_large_destroy(SapxLarge* s) {
    if (!s) return;
    sapx_chol_destroy(s->chol);
    for (void* p : {(void*)s->H, (void*)s->v, (void*)s->dsc, (void*)s->grad, (void*)s->jimp, (void*)s->Av, (void*)s->rAr,
                    (void*)s->hd, (void*)s->mcost, (void*)s->dv, (void*)s->rhs, (void*)s->part3, (void*)s->kc,
                    (void*)s->gam, (void*)s->ccost, (void*)s->vc, (void*)s->dvc, (void*)s->G, (void*)s->red})
        (void)hipFree(p);
    if (s->hred) (void)hipHostFree(s->hred);
    delete s;
}

int32_t sapx_large_solve(SapxLarge* s, const SapxLargeProblem* p, const SapxLargeOpts* o, double* v_out,
                         double* gamma_out, SapxLargeStats* stats, void* stream) {
    if (!s || !p || !o || !v_out || !stats) return lfail(SAPX_ERR_INVALID, "sapx_large_solve: null argument");
    const int n = p->n, nc = p->contact_count, K = s->kmax;
    if (n <= 0 || n > s->max_n) return lfail(SAPX_ERR_INVALID, "sapx_large_solve: n outside 1..max_dofs");
    if (nc < 0 || nc > s->max_nc) return lfail(SAPX_ERR_CAPACITY, "sapx_large_solve: more contacts than max_contacts");
    if (p->kmax != K) return lfail(SAPX_ERR_INVALID, "sapx_large_solve: kmax differs from the solver's");
    if (!p->A || !p->v_star || !p->v_guess || !p->participating ||
        (nc > 0 && (!p->J_cols || !p->J_vals || !p->phi0 || !p->weight || !p->stiffness || !p->tau || !p->mu)))
        return lfail(SAPX_ERR_INVALID, "sapx_large_solve: missing problem arrays");
    const bool has_model = p->pd_active != nullptr;
    if (has_model && (!p->pd_a || !p->pd_gain || !p->pd_limit || !p->lo_active || !p->up_active || !p->lo_vhat ||
                      !p->up_vhat || !p->lo_r || !p->up_r))
        return lfail(SAPX_ERR_INVALID, "sapx_large_solve: model terms need every array (or none)");
    if (o->line_search != SAPX_LS_EXACT_ROOT && o->line_search != SAPX_LS_ARMIJO_DECAY)
        return lfail(SAPX_ERR_UNSUPPORTED, "sapx_large_solve: line searches exact_root and armijo_decay only");
    if (!(p->dt > 0.0)) return lfail(SAPX_ERR_INVALID, "sapx_large_solve: dt must be positive");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    ModelTerms m{p->pd_active, p->pd_a, p->pd_gain, p->pd_limit, p->lo_active, p->up_active,
                 p->lo_vhat,   p->up_vhat, p->lo_r,  p->up_r,    p->dt};
    const int gn = (n + NT - 1) / NT, gc = (nc + NT - 1) / NT, grow = (n + 3) / 4;
    std::memset(stats, 0, sizeof(*stats));
    if (nc > 0)
        hipLaunchKernelGGL(l_constants, dim3(gc), dim3(NT), 0, st, nc, p->phi0, p->weight, p->stiffness, p->tau, p->mu,
                           p->dt, o->beta, o->sigma, s->kc);
    hipLaunchKernelGGL(l_init, dim3(gn), dim3(NT), 0, st, n, p->A, p->v_star, p->v_guess, p->participating, s->v,
                       s->dsc);
    // unconstrained: no contact and no active model term -> v*
    bool any_model = false;
    if (has_model && nc == 0) {
        // reading three masks is the host's only look at the problem: cheap next to a solve
        int* h = new int[3 * (size_t)n];
        bool okc = hipMemcpyAsync(h, p->pd_active, n * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipMemcpyAsync(h + n, p->lo_active, n * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipMemcpyAsync(h + 2 * n, p->up_active, n * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipStreamSynchronize(st) == hipSuccess;
        for (int i = 0; okc && i < 3 * n; ++i) any_model = any_model || h[i] != 0;
        delete[] h;
        if (!okc) return lfail(SAPX_ERR_HIP, "sapx_large_solve: reading the model-term masks failed");
    }
    if (nc == 0 && !any_model) {
        hipLaunchKernelGGL(l_copy, dim3(gn), dim3(NT), 0, st, n, p->v_star, v_out);
        stats->converged = 1;
        if (hipStreamSynchronize(st) != hipSuccess) return lfail(SAPX_ERR_HIP, "sapx_large_solve: launch failed");
        return SAPX_OK;
    }
    const double cost_min_alpha = 0.5;
    double alpha = 1.0, previous_cost = 0.0, last_cost = 0.0;
    int iteration = 0, ls_total = 0, status = 0;
    double* ls_rec = s->red + 4;
    for (;;) {
        // evaluate at v
        hipLaunchKernelGGL(l_model, dim3(gn), dim3(NT), 0, st, n, m, s->v, s->jimp, s->hd, s->mcost);
        if (nc > 0)
            hipLaunchKernelGGL(l_contacts, dim3(gc), dim3(NT), 0, st, nc, K, n, p->J_cols, p->J_vals, s->kc, s->v,
                               s->gam, s->G, s->ccost, s->jimp);
        hipLaunchKernelGGL(l_rows, dim3(grow), dim3(NT), 0, st, n, p->A, s->v, p->v_star, s->jimp, s->Av, s->grad,
                           s->rAr);
        hipLaunchKernelGGL(l_eval_reduce, dim3(1), dim3(NR), 0, st, n, nc, s->rAr, s->ccost, s->mcost, s->grad, s->Av,
                           s->jimp, s->dsc, p->participating, o->optimality_abs_tol, o->optimality_rel_tol, s->red);
        hipLaunchKernelGGL(l_m0, dim3(1), dim3(NR), 0, st, n, s->rAr, s->red);
        if (hipMemcpyAsync(s->hred, s->red, kRedSlots * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return lfail(SAPX_ERR_HIP, "sapx_large_solve: evaluation failed");
        const double cost = s->hred[0], gnorm = s->hred[1], otol = s->hred[2];
        if (iteration > 0) {
            // the previous line search: an accepted step sets alpha, the previous cost and the
            // trial count; a rejected one leaves them (contact_solve.py:6129-6285)
            status = (int)s->hred[7];
            if (s->hred[4] != 0.0) {
                alpha = s->hred[5];
                previous_cost = last_cost;
                ls_total += (int)s->hred[6];
            }
        }
        const bool opt_reached = gnorm <= otol;
        bool cost_reached = false;
        if (iteration > 0 && alpha > cost_min_alpha) {
            const double scale = 0.5 * (std::fabs(cost) + std::fabs(previous_cost));
            if (std::fabs(cost - previous_cost) < o->cost_abs_tol + o->cost_rel_tol * scale) cost_reached = true;
        }
        stats->iterations = iteration;
        stats->ls_iterations = ls_total;
        stats->cost = cost;
        stats->grad_norm = gnorm;
        stats->opt_tol = otol;
        stats->status = status;
        if (opt_reached || cost_reached) {
            stats->converged = 1;
            stats->optimality_reached = opt_reached ? 1 : 0;
            break;
        }
        if (iteration >= o->max_iterations) break;
        // Newton direction: (A + J^T G J + diag) dv = -grad
        hipLaunchKernelGGL(l_hinit, dim3(std::min<size_t>(((size_t)n * n + NT - 1) / NT, 8192)), dim3(NT), 0, st, n,
                           p->A, s->hd, o->diag_shift, s->H);
        if (nc > 0)
            hipLaunchKernelGGL(l_hcontacts, dim3(nc), dim3(64), (size_t)4 * K * sizeof(float), st, n, K, p->J_cols,
                               p->J_vals, s->G, s->H);
        int32_t rc = sapx_chol_factorize(s->chol, s->H, n, n, (int64_t)n * n, stream);
        if (rc != SAPX_OK) return rc;
        hipLaunchKernelGGL(l_neg, dim3(gn), dim3(NT), 0, st, n, s->grad, s->rhs);
        rc = sapx_chol_solve(s->chol, s->rhs, s->dv, o->refine, stream);
        if (rc != SAPX_OK) return rc;
        // line search, all trials on the device
        if (nc > 0)
            hipLaunchKernelGGL(l_dvc, dim3(gc), dim3(NT), 0, st, nc, K, n, p->J_cols, p->J_vals, s->dv, s->v, s->dvc,
                               s->vc);
        hipLaunchKernelGGL(l_rows_dv, dim3(grow), dim3(NT), 0, st, n, p->A, s->dv, s->v, p->v_star, s->grad, s->part3);
        LsArgs la;
        la.n = n;
        la.nc = nc;
        la.variant = o->line_search;
        la.max_it = o->ls_max_iterations;
        la.armijo_c = o->armijo_c;
        la.rho = o->rho;
        la.relative_slop = o->relative_slop;
        la.cost_abs_tol = o->cost_abs_tol;
        la.cost_rel_tol = o->cost_rel_tol;
        la.kc = s->kc;
        la.vc = s->vc;
        la.dvc = s->dvc;
        la.v = s->v;
        la.dv = s->dv;
        la.part3 = s->part3;
        la.red = s->red;
        la.m = m;
        la.out = ls_rec;
        hipLaunchKernelGGL(l_line_search, dim3(1), dim3(NR), 0, st, la);
        hipLaunchKernelGGL(l_update, dim3(gn), dim3(NT), 0, st, n, ls_rec, s->dv, s->v);
        last_cost = cost;
        ++iteration;
    }
    hipLaunchKernelGGL(l_copy, dim3(gn), dim3(NT), 0, st, n, s->v, v_out);
    if (gamma_out && nc > 0) hipLaunchKernelGGL(l_copy, dim3((3 * nc + NT - 1) / NT), dim3(NT), 0, st, 3 * nc, s->gam,
                                                gamma_out);
    if (hipStreamSynchronize(st) != hipSuccess) return lfail(SAPX_ERR_HIP, "sapx_large_solve: launch failed");
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lfail(SAPX_ERR_HIP, std::string("sapx_large_solve: ") + hipGetErrorString(e));
    return SAPX_OK;
}

}  // extern "C"
