// adjoint.hpp -- discrete adjoint of the colour-gradient step (periodic
// domain, single relaxation rate, FP32 populations) on the GPU, and the same
// adjoint in FP64 on the CPU as its cross-check.
//
// State and adjoint live in "node layout": f_i(x) = the population that
// arrives at node x (what the reference stores after streaming), rows as in
// the FP32 population row blocks but aligned at the receiving node.  One
// forward step is f^{n+1}_i(x + c_i) = C_i(m(f^n(x)), F(psi(f^n)) at x), so
// for the adjoint lambda of f^{n+1}:
//   mu_i(x)      = lambda_i(x + c_i)                        (streaming^T)
//   lam_m, lam_F = VJP of the collision at x                 (adj_vjp)
//   Lpsi(y)      = sum_k 3 w_k c_k . lam_F(y - c_k)          (colour gradient^T)
//   lam_m(y)    += Lpsi(y) dpsi/d(rr, rb)
//   lambda^n(y)  = VJP of the moments pass at y              (adj_moments)
// plus dJ/dA and dJ/dbeta summed over nodes in FP64.  All index shifts wrap
// periodically over the interior.
#pragma once

#include <cmath>
#include <vector>

#include "adjoint_node.hpp"
#include "kernels.hpp"

namespace mcg {

__device__ __forceinline__ float mom_psi(const DevParams &P, const float *__restrict__ mom, uint32_t o) {
    return psi(mom[o], mom[size_t(P.plane) + o]);
}

// node-layout element of direction i (colour c) at interior node (x,y,z)
__device__ __forceinline__ size_t nrow(const DevParams &P, int c, int i, int x, int y, int z) {
    return size_t(uint32_t(c * Q + i) * uint32_t(P.PX)) + pidx64(P, x, y, z);
}

// wave sum of a double (64 lanes)
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

#ifndef MCG_ADJ_WAVES_PER_EU
#define MCG_ADJ_WAVES_PER_EU 4
#endif

// psi of every interior node into one plane (the colour gradients of the
// adjoint read 26 neighbours each)
__global__ __launch_bounds__(256) void k_psi_plane(DevParams P, const float *__restrict__ mom, float *__restrict__ ps) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const uint32_t o = midx(P, x, y, z);
    ps[o] = mom_psi(P, mom, o);
}

// Jet-case inputs of the collision adjoint.
struct AdjJet {
    const float *psf;   // psi over the full grid (k_psi_full)
    const float *w_uy;  // seed on u_y (full grid) or null
};

// Adjoint modes: periodic domain, jet (inlet/outlet planes), x/z walls with
// periodic y.
enum { kAdjPeriodic = 0, kAdjJet = 1, kAdjWall = 2 };

// psi at every position of the full grid as the colour gradient reads it.
// Jet cases (forces.cuh:31-33, no wrap): interior nodes from the moments, the
// inner columns of the outlet plane NY-1 from this step's outlet state gBn,
// every other ghost position from the ghost densities.  Walls: the same with
// the y ghost planes holding the wrapped planes NY-2 and 1.
__global__ __launch_bounds__(256) void k_psi_full(DevParams P, const float *__restrict__ mom, const float *__restrict__ gBn,
                                                  const float *__restrict__ grr, const float *__restrict__ grb,
                                                  float *__restrict__ psf) {
    const int x = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int yo = int(blockIdx.y) * 4 + int(threadIdx.y);
    const int z = int(blockIdx.z);
    if (x >= P.NX || yo >= P.NY) return;
    const size_t go = gidx(P, x, yo, z);
    const int y = P.wall ? wrapN(yo, P.NY) : yo;
    const size_t g = gidx(P, x, y, z);
    const bool inner = x >= 1 && x <= P.NX - 2 && z >= 1 && z <= P.NZ - 2;
    float v;
    if (inner && y >= 1 && y <= P.NY - 2) {
        v = mom_psi(P, mom, midx(P, x, y, z));
    } else if (inner && y == P.NY - 1) {
        const size_t nxz = size_t(P.NX) * P.NZ, o = size_t(z) * P.NX + x;
        v = psi(gBn[o], gBn[nxz + o]);
    } else {
        v = psi(grr[g], grb[g]);
    }
    psf[go] = v;
}

// Adjoint of the collision at every interior node: lam_m (11 planes),
// lam_F (3 planes), and the workgroup's sums of dJ/dA, dJ/dbeta, dJ/dgravity
// in part[3 b .. 3 b + 2] (b = linear workgroup id; reduced by k_adj_sum in
// a fixed order, so the FP64 parameter gradients are deterministic).
// Jet cases: the colour gradient reads psi without wrap, ghost positions from
// the ghost densities except the outlet plane NY-1 (its inner columns hold
// the outlet state gBn of this step); populations that stream through the
// planes y=0 / y=NY-1 leave the domain; w_uy (optional) seeds u_y.  Walls:
// populations streaming into an x/z wall come back to the node in the
// opposite direction; the gradient reads the frozen ghost psi there.
template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MCG_ADJ_WAVES_PER_EU)))
void k_adj_node(DevParams P, const float *__restrict__ mom, const float *__restrict__ ps,
                const float *__restrict__ lam_in, float *__restrict__ lam_m, float *__restrict__ lam_F, double *part,
                AdjJet J) {
    __shared__ double wsum[3][4];
    const int x = 1 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int y = 1 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int z = 1 + int(blockIdx.z);
    const bool live = x <= P.NX - 2 && y <= P.NY - 2;
    double gA = 0.0, gB = 0.0, gG = 0.0;
    if (live) {
        const uint32_t o = midx(P, x, y, z);
        float m[11];
        for (int k = 0; k < 11; ++k) m[k] = mom[size_t(k) * P.plane + o];
        float Fx = 0.f, Fy = 0.f, Fz = 0.f;
        auto phi = [&](auto Ic) -> float {
            constexpr int I = decltype(Ic)::value;
            if constexpr (MODE != kAdjPeriodic) return J.psf[gidx(P, x + kCX[I], y + kCY[I], z + kCZ[I])];
            return ps[midx(P, wrapN(x + kCX[I], P.NX), wrapN(y + kCY[I], P.NY), wrapN(z + kCZ[I], P.NZ))];
        };
        color_gradient(phi, Fx, Fy, Fz);
        const float omega = omega_node(P, y, m[0], m[1]);
        AdjExt<float> ext;
        ext.iface = P.visc_mode == MCG_VISC_INTERFACE;
        ext.taur = P.taur;
        ext.taub = P.taub;
        ext.grav = P.grav;
        const AdjNode<float> n = adj_setup<float>(m, Fx, Fy, Fz, A_node(P, m[0], m[1]), P.beta, omega, ext);
        const size_t self = pidx64(P, x, y, z);
        // mu_i(x) = lambda_i(x + c_i), loaded where the pair needs it
        const AdjOut<float> a = adj_vjp_fn<float>(n, [&](int i, float &r, float &b) {
            size_t q;
            int row = i;
            if constexpr (MODE == kAdjJet) {
                const int ty = y + kCY[i];
                if (ty < 1 || ty > P.NY - 2) {  // leaves through the inlet or outlet plane
                    r = b = 0.0f;
                    return;
                }
                q = pidx64(P, wrapN(x + kCX[i], P.NX), ty, wrapN(z + kCZ[i], P.NZ));
            } else if constexpr (MODE == kAdjWall) {
                const int tx = x + kCX[i], tz = z + kCZ[i];
                if (tx < 1 || tx > P.NX - 2 || tz < 1 || tz > P.NZ - 2) {  // bounced back into direction opp(i)
                    q = self;
                    row = opp(i);
                } else {
                    q = pidx64(P, tx, wrapN(y + kCY[i], P.NY), tz);
                }
            } else {
                q = pidx64(P, wrapN(x + kCX[i], P.NX), wrapN(y + kCY[i], P.NY), wrapN(z + kCZ[i], P.NZ));
            }
            r = lam_in[size_t(uint32_t(row) * uint32_t(P.PX)) + q];
            b = lam_in[size_t(uint32_t(Q + row) * uint32_t(P.PX)) + q];
        });
        float lm[11];
        for (int k = 0; k < 11; ++k) lm[k] = a.m[k];
        if (MODE == kAdjJet && J.w_uy) lm[3] += J.w_uy[gidx(P, x, y, z)];
        for (int k = 0; k < 11; ++k) lam_m[size_t(k) * P.plane + o] = lm[k];
        for (int k = 0; k < 3; ++k) lam_F[size_t(k) * P.plane + o] = a.F[k];
        gA = double(a.A);
        gB = double(a.beta);
        gG = double(a.grav);
        if (P.grav != 0.0f) {
            // the moments pass: u_y = (j_y - grav rr/2)/rho; the adjoint of u_y includes Pi's share
            const float rho = m[0] + m[1];
            const float luy = lm[3] - 2.0f * rho * m[3] * lm[5 + PYY] - rho * (m[2] * lm[5 + PXY] + m[4] * lm[5 + PYZ]);
            gG += -0.5 * double(m[0]) / double(rho) * double(luy);
        }
    }
    gA = wave_sum(gA);
    gB = wave_sum(gB);
    gG = wave_sum(gG);
    if (threadIdx.x == 0) {
        wsum[0][threadIdx.y] = gA;
        wsum[1][threadIdx.y] = gB;
        wsum[2][threadIdx.y] = gG;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const size_t b = size_t(blockIdx.x) + size_t(gridDim.x) * (size_t(blockIdx.y) + size_t(gridDim.y) * blockIdx.z);
        for (int v = 0; v < 3; ++v) part[3 * b + v] = (wsum[v][0] + wsum[v][1]) + (wsum[v][2] + wsum[v][3]);
    }
}

// acc[0..NV-1] += the sums of the nb workgroup partials part[NV b + v] (one
// workgroup of 1024 threads, four independent accumulators each, fixed
// order: deterministic)
constexpr int kSumThreads = 1024;
template <int NV>
__global__ __launch_bounds__(kSumThreads) void k_adj_sum(const double *__restrict__ part, size_t nb, double *acc) {
    __shared__ double sh[NV][kSumThreads];
    double a[NV][4];
    for (int v = 0; v < NV; ++v) a[v][0] = a[v][1] = a[v][2] = a[v][3] = 0.0;
    size_t k = threadIdx.x;
    for (; k + 3 * kSumThreads < nb; k += 4 * kSumThreads) {
        for (int u = 0; u < 4; ++u)
            for (int v = 0; v < NV; ++v) a[v][u] += part[NV * (k + u * kSumThreads) + v];
    }
    for (; k < nb; k += kSumThreads)
        for (int v = 0; v < NV; ++v) a[v][0] += part[NV * k + v];
    for (int v = 0; v < NV; ++v) sh[v][threadIdx.x] = (a[v][0] + a[v][1]) + (a[v][2] + a[v][3]);
    __syncthreads();
    for (int h = kSumThreads / 2; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h)
            for (int v = 0; v < NV; ++v) sh[v][threadIdx.x] += sh[v][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < NV; ++v) acc[v] += sh[v][0];
}

// Per-workgroup partial sums of w(x) u_y(x) over the interior nodes (the
// time-windowed u_y functional), part[2 b] (part[2 b + 1] = 0), for k_adj_sum.
__global__ __launch_bounds__(256) void k_uy_partial(DevParams P, const float *__restrict__ mom, const float *__restrict__ w,
                                                    double *part) {
    const int x = 1 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int y = 1 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int z = 1 + int(blockIdx.z);
    double v = 0.0;
    if (x <= P.NX - 2 && y <= P.NY - 2) v = double(w[gidx(P, x, y, z)]) * double(mom[3 * size_t(P.plane) + midx(P, x, y, z)]);
    __shared__ double sh[4];
    v = wave_sum(v);
    const int wv = int(threadIdx.y);
    if (threadIdx.x == 0) sh[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const size_t b = (size_t(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        part[2 * b] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        part[2 * b + 1] = 0.0;
    }
}

// Colour-gradient transpose, psi derivative and moments-pass adjoint:
// lambda^n at every interior node.  Jet cases: the transpose runs over the
// interior neighbours without wrap, and the stale inlet/outlet slots (never
// rewritten, so carried unchanged) pass their adjoint through.
template <int MODE>
__global__ __launch_bounds__(256) void k_adj_pop(DevParams P, const float *__restrict__ mom,
                                                 const float *__restrict__ lam_m, const float *__restrict__ lam_F,
                                                 const float *__restrict__ lam_in, float *__restrict__ lam_out) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const uint32_t o = midx(P, x, y, z);
    float Lpsi = 0.f;
    unroll<1, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        int sx = x - kCX[I], sy = y - kCY[I], sz = z - kCZ[I];
        if constexpr (MODE == kAdjJet) {
            if (sx < 1 || sx > P.NX - 2 || sy < 1 || sy > P.NY - 2 || sz < 1 || sz > P.NZ - 2) return;
        } else if constexpr (MODE == kAdjWall) {
            if (sx < 1 || sx > P.NX - 2 || sz < 1 || sz > P.NZ - 2) return;  // ghost psi is constant
            sy = wrapN(sy, P.NY);
        } else {
            sx = wrapN(sx, P.NX);
            sy = wrapN(sy, P.NY);
            sz = wrapN(sz, P.NZ);
        }
        const uint32_t q = midx(P, sx, sy, sz);
        float d = 0.f;
        if constexpr (kCX[I] != 0) d += float(kCX[I]) * lam_F[q];
        if constexpr (kCY[I] != 0) d += float(kCY[I]) * lam_F[size_t(P.plane) + q];
        if constexpr (kCZ[I] != 0) d += float(kCZ[I]) * lam_F[2 * size_t(P.plane) + q];
        Lpsi += inv_cs2 * w(I) * d;
    });
    float m[11], lam[11];
    for (int k = 0; k < 11; ++k) {
        m[k] = mom[size_t(k) * P.plane + o];
        lam[k] = lam_m[size_t(k) * P.plane + o];
    }
    const float rho = m[0] + m[1];
    const float inv2 = 1.0f / (rho * rho);
    lam[0] += Lpsi * 2.0f * m[1] * inv2;
    lam[1] += -Lpsi * 2.0f * m[0] * inv2;
    float lr[Q], lb[Q];
    adj_moments<float>(m, lam, lr, lb, P.ghalf);
    if constexpr (MODE == kAdjJet) {
        auto inner = [&](int a, int c) { return a >= 1 && a <= P.NX - 2 && c >= 1 && c <= P.NZ - 2; };
        for (int i = 0; i < Q; ++i) {
            const int gx = x - kCX[i], gz = z - kCZ[i];
            const bool stale = (y == 1 && kCY[i] == 1 && !(inner(gx, gz) && is_jet(P, gx, gz))) ||
                               (y == P.NY - 2 && kCY[i] == -1 && !inner(gx, gz));
            if (stale) {
                lr[i] += lam_in[nrow(P, 0, i, x, y, z)];
                lb[i] += lam_in[nrow(P, 1, i, x, y, z)];
            }
        }
    }
    for (int i = 0; i < Q; ++i) {
        __builtin_nontemporal_store(lr[i], &lam_out[nrow(P, 0, i, x, y, z)]);
        __builtin_nontemporal_store(lb[i], &lam_out[nrow(P, 1, i, x, y, z)]);
    }
}

// Inlet velocity of jet column (x,z) at step n (inlet_velocity) with its
// derivatives with respect to jet_velocity (dU) and perturbation_amp (damp).
__device__ __forceinline__ void inlet_velocity_grad(const DevParams &P, const float *du, uint64_t step, int x, int z,
                                                    float e[3], float dU[3], float damp[3]) {
    inlet_velocity(P, du, step, x, z, e[0], e[1], e[2]);
    dU[0] = dU[2] = damp[0] = damp[1] = damp[2] = 0.0f;
    dU[1] = 1.0f;
    if (P.inlet_profile == MCG_INLET_TANH) {
        const float dx = float(x) - P.jet_x0, dz = float(z) - P.jet_z0;
        dU[1] = 0.5f * (1.0f - tanhf((sqrtf(dx * dx + dz * dz) - P.jet_R) / P.inlet_w));
        const float sx = 2.0f * inlet_xi(P.pert_seed, step, x, z, P.NX, 1) - 1.0f;
        const float sz = 2.0f * inlet_xi(P.pert_seed, step, x, z, P.NX, 2) - 1.0f;
        dU[0] = P.pert_amp * sx;
        dU[2] = P.pert_amp * sz;
        damp[0] = P.jet_U * sx;
        damp[2] = P.jet_U * sz;
    }
}

// block sum of two doubles into part[2 b], part[2 b + 1] (256 threads)
__device__ __forceinline__ void block_part2(double a, double b, double *part) {
    __shared__ double sh[2][4];
    a = wave_sum(a);
    b = wave_sum(b);
    const int wv = int(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][wv] = a;
        sh[1][wv] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        part[2 * blockIdx.x + 1] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

// Boundary adjoint of a jet step, one thread per inner column (x,z); runs
// after k_adj_node (reads lam_F, adds to lam_m at (x,1,z) and (x,NY-2,z),
// which only this thread touches).
//  * convective outlet (jet/boundary_conditions.cuh:163-233): the 9 emitted
//    populations and the update of the outlet state; lgB_out = adjoint of
//    this step's outlet state gBn, from lgB_in = adjoint of the next one;
//  * psi of the outlet ghost plane, read by the colour gradient of row NY-2:
//    its transpose, into lgB_out;
//  * inlet of a jet column (jet/boundary_conditions.cuh:17-85): adds to
//    dJ/de of the column (FP64, summed over steps) and the dU, damp partials
//    part[2 b], part[2 b + 1].
__global__ __launch_bounds__(256) void k_adj_cols(DevParams P, const float *__restrict__ mom,
                                                  const float *__restrict__ lam_in, const float *__restrict__ lam_F,
                                                  float *__restrict__ lam_m, const float *__restrict__ gBn,
                                                  const float *__restrict__ lgB_in, float *__restrict__ lgB_out,
                                                  double *__restrict__ dinlet, double *part, unsigned long long step,
                                                  const float *__restrict__ inlet_du) {
    const int c = int(blockIdx.x) * 256 + int(threadIdx.x);
    const int nx = P.NX - 2;
    const int x = 1 + c % nx, z = 1 + c / nx;
    double pU = 0.0, pA = 0.0;
    if (z <= P.NZ - 2) {
        const size_t nxz = size_t(P.NX) * P.NZ, oc = size_t(z) * P.NX + x;
        // outlet
        {
            const uint32_t o = midx(P, x, P.NY - 2, z);
            float m[11];
            for (int k = 0; k < 11; ++k) m[k] = mom[size_t(k) * P.plane + o];
            const float uc = clamp01(m[3]);
            float gB[5], g1[5];
            for (int k = 0; k < 5; ++k) {
                gB[k] = gBn[k * nxz + oc];
                g1[k] = gB[k] - uc * (gB[k] - m[k]);
            }
            const float rhoT = g1[0] + g1[1];
            const float inv = 1.0f / rhoT;
            const float aR = g1[0] * inv, aB = g1[1] * inv;
            const float oms = 1.0f - omega_node(P, P.NY - 2, g1[0], g1[1]);
            float l_aR = 0.f, l_aB = 0.f, l_rhoT = 0.f, l_uB[3] = {0.f, 0.f, 0.f}, lP[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f},
                  lu[3] = {0.f, 0.f, 0.f};
            for (int k = 0; k < 9; ++k) {
                const int i = kDown[k], tx = x + kCX[i], tz = z + kCZ[i];
                if (tx < 1 || tx > P.NX - 2 || tz < 1 || tz > P.NZ - 2) continue;
                const float lrk = lam_in[nrow(P, 0, i, tx, P.NY - 2, tz)], lbk = lam_in[nrow(P, 1, i, tx, P.NY - 2, tz)];
                const float g = feqT<float>(i, rhoT, g1[2], g1[3], g1[4]) + oms * fneqrT<float>(i, m + 5, m[2], m[3], m[4]);
                l_aR += g * lrk;
                l_aB += g * lbk;
                const float lg = aR * lrk + aB * lbk;
                feq_vjp<float>(i, rhoT, g1[2], g1[3], g1[4], lg, l_rhoT, l_uB);
                fneqr_vjp<float>(i, m + 5, m[2], m[3], m[4], oms * lg, lP, lu);
            }
            const float q = (g1[0] * l_aR + g1[1] * l_aB) * inv * inv;
            float L[5];
            L[0] = l_aR * inv + l_rhoT - q + lgB_in[oc];
            L[1] = l_aB * inv + l_rhoT - q + lgB_in[nxz + oc];
            for (int k = 0; k < 3; ++k) L[2 + k] = l_uB[k] + lgB_in[(2 + k) * nxz + oc];
            float l_uc = 0.f, lo[5];
            for (int k = 0; k < 5; ++k) {
                lo[k] = (1.0f - uc) * L[k];
                lam_m[size_t(k) * P.plane + o] += uc * L[k];
                l_uc += -(gB[k] - m[k]) * L[k];
            }
            if (m[3] > 0.0f && m[3] < 1.0f) lam_m[3 * size_t(P.plane) + o] += l_uc;
            for (int k = 0; k < 6; ++k) lam_m[size_t(5 + k) * P.plane + o] += lP[k];
            for (int k = 0; k < 3; ++k) lam_m[size_t(2 + k) * P.plane + o] += lu[k];
            // outlet ghost psi (c_y = +1 neighbours of row NY-2 read it)
            float Lpsi = 0.f;
            unroll<1, Q>([&](auto Ic) {
                constexpr int I = decltype(Ic)::value;
                if constexpr (kCY[I] == 1) {
                    const int sx = x - kCX[I], sz = z - kCZ[I];
                    if (sx < 1 || sx > P.NX - 2 || sz < 1 || sz > P.NZ - 2) return;
                    const uint32_t qn = midx(P, sx, P.NY - 2, sz);
                    float d = lam_F[size_t(P.plane) + qn];
                    if constexpr (kCX[I] != 0) d += float(kCX[I]) * lam_F[qn];
                    if constexpr (kCZ[I] != 0) d += float(kCZ[I]) * lam_F[2 * size_t(P.plane) + qn];
                    Lpsi += inv_cs2 * w(I) * d;
                }
            });
            const float rho = gB[0] + gB[1];
            lo[0] += Lpsi * 2.0f * gB[1] / (rho * rho);
            lo[1] += -Lpsi * 2.0f * gB[0] / (rho * rho);
            for (int k = 0; k < 5; ++k) lgB_out[k * nxz + oc] = lo[k];
        }
        // inlet
        if (is_jet(P, x, z)) {
            const uint32_t o = midx(P, x, 1, z);
            float m[11];
            for (int k = 0; k < 11; ++k) m[k] = mom[size_t(k) * P.plane + o];
            // jet node: red share 0, blue share 1, total density rhob0
            const float rT = P.rhob0;
            const float oms = 1.0f - P.omega_b;
            float e[3], dU[3], da[3], le[3] = {0.f, 0.f, 0.f}, lP[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, lu[3] = {0.f, 0.f, 0.f};
            inlet_velocity_grad(P, inlet_du, step, x, z, e, dU, da);
            for (int k = 0; k < 9; ++k) {
                const int i = kUp[k], tx = x + kCX[i], tz = z + kCZ[i];
                if (tx < 1 || tx > P.NX - 2 || tz < 1 || tz > P.NZ - 2) continue;
                const float lg = lam_in[nrow(P, 1, i, tx, 1, tz)];
                float l_rho = 0.f;
                feq_vjp<float>(i, rT, e[0], e[1], e[2], lg, l_rho, le);
                fneqr_vjp<float>(i, m + 5, m[2], m[3], m[4], oms * lg, lP, lu);
            }
            for (int k = 0; k < 6; ++k) lam_m[size_t(5 + k) * P.plane + o] += lP[k];
            for (int k = 0; k < 3; ++k) lam_m[size_t(2 + k) * P.plane + o] += lu[k];
            for (int k = 0; k < 3; ++k) {
                dinlet[k * nxz + oc] += double(le[k]);
                pU += double(le[k]) * double(dU[k]);
                pA += double(le[k]) * double(da[k]);
            }
        }
    }
    block_part2(pU, pA, part);
}

// reference layout (i*Ncells + id, ghosts included) <-> node layout rows
__global__ __launch_bounds__(256) void k_rows_from_ref(DevParams P, const float *__restrict__ ref_r,
                                                       const float *__restrict__ ref_b, float *__restrict__ rows) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const size_t N = size_t(P.NX) * P.NY * P.NZ;
    const size_t g = gidx(P, x, y, z);
    for (int i = 0; i < Q; ++i) {
        rows[nrow(P, 0, i, x, y, z)] = ref_r[size_t(i) * N + g];
        rows[nrow(P, 1, i, x, y, z)] = ref_b[size_t(i) * N + g];
    }
}

__global__ __launch_bounds__(256) void k_rows_to_ref(DevParams P, const float *__restrict__ rows, float *__restrict__ ref_r,
                                                     float *__restrict__ ref_b) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const size_t N = size_t(P.NX) * P.NY * P.NZ;
    const size_t g = gidx(P, x, y, z);
    for (int i = 0; i < Q; ++i) {
        ref_r[size_t(i) * N + g] = rows[nrow(P, 0, i, x, y, z)];
        ref_b[size_t(i) * N + g] = rows[nrow(P, 1, i, x, y, z)];
    }
}

// ------------------------------------------------------------------ CPU FP64
// The same forward and adjoint on the host in FP64 over reference-layout
// arrays (periodic interior, ghosts unused).  It serves as the FP64 adjoint
// that the GPU adjoint is compared with, and it is itself checked against
// central finite differences of the independent FP64 oracle step.
struct CpuAdjoint {
    int NX, NY, NZ;
    double A, beta, omega;
    size_t N;
    // extensions: x/z walls (bounce-back, ghost psi frozen at gpsi), per-node
    // omega and the body force on red
    bool wall = false;
    AdjExt<double> ext{};
    std::vector<double> gpsi;  // psi of every position of the initial state (walls)

    size_t id(int x, int y, int z) const { return size_t(x) + size_t(NX) * (size_t(y) + size_t(NY) * size_t(z)); }
    static int wrap(int v, int n) { return v == 0 ? n - 2 : (v == n - 1 ? 1 : v); }
    bool inner_xz(int x, int z) const { return x >= 1 && x <= NX - 2 && z >= 1 && z <= NZ - 2; }
    // streaming target of direction i from node (x,y,z); false: bounced back at a wall
    bool target(int x, int y, int z, int i, size_t &t) const {
        int tx = x + kCX[i], tz = z + kCZ[i];
        if (wall) {
            if (!inner_xz(tx, tz)) return false;
        } else {
            tx = wrap(tx, NX);
            tz = wrap(tz, NZ);
        }
        t = id(tx, wrap(y + kCY[i], NY), tz);
        return true;
    }

    void moments(const double *fr, const double *fb, int x, int y, int z, double m[11]) const {
        const size_t g = id(x, y, z);
        double sr = 0, sb = 0, j[3] = {0, 0, 0}, S[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < Q; ++i) {
            const double a = fr[size_t(i) * N + g], b = fb[size_t(i) * N + g], gg = a + b;
            sr += a;
            sb += b;
            j[0] += kCX[i] * gg;
            j[1] += kCY[i] * gg;
            j[2] += kCZ[i] * gg;
            for (int k = 0; k < 6; ++k) S[k] += gg * h2T<double>(k, i);
        }
        const double rho = sr + sb;
        m[0] = sr;
        m[1] = sb;
        j[1] -= 0.5 * ext.grav * sr;  // u = (j + F/2)/rho, F_y = -grav rho_r
        for (int a = 0; a < 3; ++a) m[2 + a] = j[a] / rho;
        const double *u = m + 2;
        m[5 + PXX] = S[PXX] - rho * u[0] * u[0];
        m[5 + PXY] = S[PXY] - rho * u[0] * u[1];
        m[5 + PYY] = S[PYY] - rho * u[1] * u[1];
        m[5 + PYZ] = S[PYZ] - rho * u[1] * u[2];
        m[5 + PZZ] = S[PZZ] - rho * u[2] * u[2];
        m[5 + PXZ] = S[PXZ] - rho * u[0] * u[2];
    }

    // moments of every interior node (11 arrays of N) and psi of every position
    void all_moments(const double *fr, const double *fb, std::vector<double> &M, std::vector<double> &ps) const {
        M.assign(11 * N, 0.0);
        if (wall) {
            ps = gpsi;
        } else {
            ps.assign(N, 0.0);
        }
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    double m[11];
                    moments(fr, fb, x, y, z, m);
                    const size_t g = id(x, y, z);
                    for (int k = 0; k < 11; ++k) M[size_t(k) * N + g] = m[k];
                    ps[g] = (m[0] - m[1]) / (m[0] + m[1]);
                }
    }

    // psi read by the colour gradient at neighbour (x,y,z): y wrapped; x/z
    // wrapped (periodic) or ghost values (walls)
    double psi_at(const std::vector<double> &ps, int x, int y, int z) const {
        if (!wall) return ps[id(wrap(x, NX), wrap(y, NY), wrap(z, NZ))];
        return ps[id(x, wrap(y, NY), z)];
    }

    AdjNode<double> node(const std::vector<double> &M, const std::vector<double> &ps, int x, int y, int z) const {
        const size_t g = id(x, y, z);
        double m[11];
        for (int k = 0; k < 11; ++k) m[k] = M[size_t(k) * N + g];
        double F[3] = {0, 0, 0};
        for (int i = 1; i < Q; ++i) {
            const double p = psi_at(ps, x + kCX[i], y + kCY[i], z + kCZ[i]);
            F[0] += 3.0 * wT<double>(i) * p * kCX[i];
            F[1] += 3.0 * wT<double>(i) * p * kCY[i];
            F[2] += 3.0 * wT<double>(i) * p * kCZ[i];
        }
        return adj_setup<double>(m, F[0], F[1], F[2], A, beta, omega, ext);
    }

    void step(const double *fr, const double *fb, double *gr, double *gb) const {
        std::vector<double> M, ps;
        all_moments(fr, fb, M, ps);
        std::fill(gr, gr + size_t(Q) * N, 0.0);
        std::fill(gb, gb + size_t(Q) * N, 0.0);
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const AdjNode<double> n = node(M, ps, x, y, z);
                    double r[Q], b[Q];
                    adj_collide<double>(n, r, b);
                    for (int i = 0; i < Q; ++i) {
                        size_t t;
                        const size_t s = target(x, y, z, i, t) ? size_t(i) * N + t : size_t(opp(i)) * N + id(x, y, z);
                        gr[s] = r[i];
                        gb[s] = b[i];
                    }
                }
    }

    // one adjoint step: lam (adjoint of f^{n+1}) -> out (adjoint of f^n)
    void adjoint(const double *fr, const double *fb, const double *lr, const double *lb, double *outr, double *outb,
                 double &dA, double &dbeta, double &dgrav) const {
        std::vector<double> M, ps;
        all_moments(fr, fb, M, ps);
        std::vector<double> LM(11 * N, 0.0), LF(3 * N, 0.0);
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const AdjNode<double> n = node(M, ps, x, y, z);
                    double mr[Q], mb[Q];
                    for (int i = 0; i < Q; ++i) {
                        size_t t;
                        const size_t s = target(x, y, z, i, t) ? size_t(i) * N + t : size_t(opp(i)) * N + id(x, y, z);
                        mr[i] = lr[s];
                        mb[i] = lb[s];
                    }
                    const AdjOut<double> a = adj_vjp<double>(n, mr, mb);
                    const size_t g = id(x, y, z);
                    for (int k = 0; k < 11; ++k) LM[size_t(k) * N + g] = a.m[k];
                    for (int k = 0; k < 3; ++k) LF[size_t(k) * N + g] = a.F[k];
                    dA += a.A;
                    dbeta += a.beta;
                    dgrav += a.grav;
                }
        std::fill(outr, outr + size_t(Q) * N, 0.0);
        std::fill(outb, outb + size_t(Q) * N, 0.0);
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const size_t g = id(x, y, z);
                    double Lpsi = 0;
                    for (int i = 1; i < Q; ++i) {
                        int sx = x - kCX[i], sz = z - kCZ[i];
                        if (wall) {
                            if (!inner_xz(sx, sz)) continue;  // ghost psi is constant
                        } else {
                            sx = wrap(sx, NX);
                            sz = wrap(sz, NZ);
                        }
                        const size_t q = id(sx, wrap(y - kCY[i], NY), sz);
                        Lpsi += 3.0 * wT<double>(i) *
                                (kCX[i] * LF[q] + kCY[i] * LF[N + q] + kCZ[i] * LF[2 * N + q]);
                    }
                    double m[11], lam[11];
                    for (int k = 0; k < 11; ++k) {
                        m[k] = M[size_t(k) * N + g];
                        lam[k] = LM[size_t(k) * N + g];
                    }
                    const double rho = m[0] + m[1];
                    lam[0] += Lpsi * 2.0 * m[1] / (rho * rho);
                    lam[1] += -Lpsi * 2.0 * m[0] / (rho * rho);
                    double r[Q], b[Q];
                    adj_moments<double>(m, lam, r, b, 0.5 * ext.grav, &dgrav);
                    for (int i = 0; i < Q; ++i) {
                        outr[size_t(i) * N + g] = r[i];
                        outb[size_t(i) * N + g] = b[i];
                    }
                }
    }
};

// ---------------------------------------------------------------- CPU FP64, jet
// Jet cases (y_mode jet, reference gradient without wrap): x/z periodic
// streaming; plane y=1 receives c_y=+1 populations from the inlet (jet
// columns) or keeps them (stale slots), plane NY-2 receives c_y=-1
// populations from the convective outlet (inner columns) or keeps them; the
// outlet state gB (rho_r, rho_b, u at y=NY-1) is carried from step to step and
// gives psi of the ghost plane NY-1 to the colour gradient; other ghost psi
// are constant.  Inlet velocity e(x,z,n) = U profile(r) + amp U (2 xi - 1)
// (sharp profile: (0, U, 0)).  The adjoint returns dJ/df at the initial state,
// dJ/dA, dJ/dbeta, dJ/dU, dJ/d(amp) and dJ/de summed over steps per jet column;
// an optional seed w(x) on u_y at the moments passes of steps first..last adds
// sum_n sum_x w(x) u_y^n(x) to J.
inline uint64_t splitmix64_host(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct CpuJetAdjoint {
    mcg_case c{};
    int NX = 0, NY = 0, NZ = 0;
    size_t N = 0;
    double A = 0, beta = 0, omega = 0;
    std::vector<double> gpsi;   // psi of every position at the initial state (ghosts constant)
    std::vector<double> du;     // optional inlet velocity offset per column (3 x NX*NZ)

    size_t id(int x, int y, int z) const { return size_t(x) + size_t(NX) * (size_t(y) + size_t(NY) * size_t(z)); }
    static int wrap(int v, int n) { return v == 0 ? n - 2 : (v == n - 1 ? 1 : v); }
    bool interior(int x, int y, int z) const { return x >= 1 && x <= NX - 2 && y >= 1 && y <= NY - 2 && z >= 1 && z <= NZ - 2; }
    bool inner(int x, int z) const { return x >= 1 && x <= NX - 2 && z >= 1 && z <= NZ - 2; }
    bool is_jet(int x, int z) const {
        const double dx = double(x) - c.jet_x0, dz = double(z) - c.jet_z0;
        return dx * dx + dz * dz <= c.jet_radius * c.jet_radius;
    }
    double xi(int64_t n, int x, int z, int k) const {
        const uint64_t key = splitmix64_host(splitmix64_host(splitmix64_host(uint64_t(uint32_t(c.perturbation_seed))) ^ uint64_t(n)) ^
                                             (uint64_t(z) * uint64_t(NX) + uint64_t(x)));
        return double(splitmix64_host(key + uint64_t(k)) >> 40) / 16777216.0;
    }
    // inlet velocity and its derivatives with respect to U and amp
    void inlet(int64_t n, int x, int z, double e[3], double dU[3], double damp[3]) const {
        double prof = 1.0;
        e[0] = e[2] = 0.0;
        dU[0] = dU[2] = damp[0] = damp[1] = damp[2] = 0.0;
        if (c.inlet_profile == MCG_INLET_TANH) {
            const double dx = double(x) - c.jet_x0, dz = double(z) - c.jet_z0;
            prof = 0.5 * (1.0 - std::tanh((std::sqrt(dx * dx + dz * dz) - c.jet_radius) / c.interface_width));
            const double sx = 2.0 * xi(n, x, z, 1) - 1.0, sz = 2.0 * xi(n, x, z, 2) - 1.0;
            e[0] = c.perturbation_amp * c.jet_velocity * sx;
            e[2] = c.perturbation_amp * c.jet_velocity * sz;
            dU[0] = c.perturbation_amp * sx;
            dU[2] = c.perturbation_amp * sz;
            damp[0] = c.jet_velocity * sx;
            damp[2] = c.jet_velocity * sz;
        }
        e[1] = c.jet_velocity * prof;
        dU[1] = prof;
        if (!du.empty()) {
            const size_t nxz = size_t(NX) * NZ, o = size_t(z) * NX + x;
            for (int k = 0; k < 3; ++k) e[k] += du[k * nxz + o];
        }
    }

    struct Step {
        std::vector<double> M, ps;  // interior moments (11 x N), psi of every position
    };

    void prepare(const double *f_r, const double *f_b, const double *gB, Step &st) const {
        st.M.assign(11 * N, 0.0);
        st.ps = gpsi;
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const size_t g = id(x, y, z);
                    double sr = 0, sb = 0, j[3] = {0, 0, 0}, S[6] = {0, 0, 0, 0, 0, 0};
                    for (int i = 0; i < Q; ++i) {
                        const double a = f_r[size_t(i) * N + g], b = f_b[size_t(i) * N + g], gg = a + b;
                        sr += a;
                        sb += b;
                        j[0] += kCX[i] * gg;
                        j[1] += kCY[i] * gg;
                        j[2] += kCZ[i] * gg;
                        for (int k = 0; k < 6; ++k) S[k] += gg * h2T<double>(k, i);
                    }
                    const double rho = sr + sb, u[3] = {j[0] / rho, j[1] / rho, j[2] / rho};
                    double *m[11];
                    for (int k = 0; k < 11; ++k) m[k] = &st.M[size_t(k) * N + g];
                    *m[0] = sr;
                    *m[1] = sb;
                    *m[2] = u[0];
                    *m[3] = u[1];
                    *m[4] = u[2];
                    *m[5 + PXX] = S[PXX] - rho * u[0] * u[0];
                    *m[5 + PXY] = S[PXY] - rho * u[0] * u[1];
                    *m[5 + PYY] = S[PYY] - rho * u[1] * u[1];
                    *m[5 + PYZ] = S[PYZ] - rho * u[1] * u[2];
                    *m[5 + PZZ] = S[PZZ] - rho * u[2] * u[2];
                    *m[5 + PXZ] = S[PXZ] - rho * u[0] * u[2];
                    st.ps[g] = (sr - sb) / rho;
                }
        // the outlet ghost plane holds the convective state of the previous step
        for (int z = 1; z < NZ - 1; ++z)
            for (int x = 1; x < NX - 1; ++x) {
                const size_t o = size_t(z) * NX + x;
                const double rr = gB[o], rb = gB[size_t(NX) * NZ + o];
                st.ps[id(x, NY - 1, z)] = (rr - rb) / (rr + rb);
            }
    }

    void mom(const Step &st, size_t g, double m[11]) const {
        for (int k = 0; k < 11; ++k) m[k] = st.M[size_t(k) * N + g];
    }

    AdjNode<double> node(const Step &st, int x, int y, int z) const {
        double m[11], F[3] = {0, 0, 0};
        mom(st, id(x, y, z), m);
        for (int i = 1; i < Q; ++i) {
            const double p = st.ps[id(x + kCX[i], y + kCY[i], z + kCZ[i])];  // no wrap (forces.cuh:31-33)
            F[0] += 3.0 * wT<double>(i) * p * kCX[i];
            F[1] += 3.0 * wT<double>(i) * p * kCY[i];
            F[2] += 3.0 * wT<double>(i) * p * kCZ[i];
        }
        return adj_setup<double>(m, F[0], F[1], F[2], A, beta, omega);
    }

    // boundary populations: inlet (9 c_y=+1 values, blue) of jet column (x,z)
    void inlet_pops(const Step &st, int64_t n, int x, int z, double b[9]) const {
        double m[11], e[3], dU[3], da[3];
        mom(st, id(x, 1, z), m);
        inlet(n, x, z, e, dU, da);
        const double oms = 1.0 - 1.0 / c.taub;
        for (int k = 0; k < 9; ++k) {
            const int i = kUp[k];
            b[k] = feqT<double>(i, c.rhob0, e[0], e[1], e[2]) + oms * fneqrT<double>(i, m + 5, m[2], m[3], m[4]);
        }
    }

    // outlet of inner column (x,z): new state gB1 and the 9 c_y=-1 values
    void outlet_pops(const Step &st, const double *gB, int x, int z, double gB1[5], double r[9], double b[9]) const {
        double m[11];
        mom(st, id(x, NY - 2, z), m);
        const size_t o = size_t(z) * NX + x, nxz = size_t(NX) * NZ;
        const double uc = std::min(std::max(m[3], 0.0), 1.0);
        const double mf[5] = {m[0], m[1], m[2], m[3], m[4]};
        for (int k = 0; k < 5; ++k) gB1[k] = gB[k * nxz + o] - uc * (gB[k * nxz + o] - mf[k]);
        const double rhoT = gB1[0] + gB1[1], aR = gB1[0] / rhoT, aB = gB1[1] / rhoT;
        const double oms = 1.0 - omega;
        for (int k = 0; k < 9; ++k) {
            const int i = kDown[k];
            const double g = feqT<double>(i, rhoT, gB1[2], gB1[3], gB1[4]) + oms * fneqrT<double>(i, m + 5, m[2], m[3], m[4]);
            r[k] = aR * g;
            b[k] = aB * g;
        }
    }

    void step(int64_t n, const double *fr, const double *fb, const double *gB, double *gr, double *gb, double *gB_out) const {
        Step st;
        prepare(fr, fb, gB, st);
        const size_t nxz = size_t(NX) * NZ;
        std::copy(fr, fr + size_t(Q) * N, gr);  // stale slots keep their values
        std::copy(fb, fb + size_t(Q) * N, gb);
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const AdjNode<double> nd = node(st, x, y, z);
                    double r[Q], b[Q];
                    adj_collide<double>(nd, r, b);
                    for (int i = 0; i < Q; ++i) {
                        const int ty = y + kCY[i];
                        if (ty < 1 || ty > NY - 2) continue;  // leaves through the inlet/outlet planes
                        const size_t t = id(wrap(x + kCX[i], NX), ty, wrap(z + kCZ[i], NZ));
                        gr[size_t(i) * N + t] = r[i];
                        gb[size_t(i) * N + t] = b[i];
                    }
                }
        std::copy(gB, gB + 5 * nxz, gB_out);
        for (int z = 1; z < NZ - 1; ++z)
            for (int x = 1; x < NX - 1; ++x) {
                if (is_jet(x, z)) {
                    double b[9];
                    inlet_pops(st, n, x, z, b);
                    for (int k = 0; k < 9; ++k) {
                        const int i = kUp[k], tx = x + kCX[i], tz = z + kCZ[i];
                        if (!inner(tx, tz)) continue;
                        const size_t t = id(tx, 1, tz);
                        gr[size_t(i) * N + t] = 0.0 * b[k];
                        gb[size_t(i) * N + t] = b[k];
                    }
                }
                double gB1[5], r[9], b[9];
                outlet_pops(st, gB, x, z, gB1, r, b);
                const size_t o = size_t(z) * NX + x;
                for (int k = 0; k < 5; ++k) gB_out[k * nxz + o] = gB1[k];
                for (int k = 0; k < 9; ++k) {
                    const int i = kDown[k], tx = x + kCX[i], tz = z + kCZ[i];
                    if (!inner(tx, tz)) continue;
                    const size_t t = id(tx, NY - 2, tz);
                    gr[size_t(i) * N + t] = r[k];
                    gb[size_t(i) * N + t] = b[k];
                }
            }
    }

    // adjoint of step n: (lr, lb, lgB) of the state after the step -> (outr, outb, outgB) before it
    void adjoint(int64_t n, const double *fr, const double *fb, const double *gB, const double *lr, const double *lb,
                 const double *lgB, double *outr, double *outb, double *outgB, double *dp, double *dinlet,
                 const double *w_uy) const {
        Step st;
        prepare(fr, fb, gB, st);
        const size_t nxz = size_t(NX) * NZ;
        std::vector<double> LM(11 * N, 0.0), LF(3 * N, 0.0);
        // collision
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const AdjNode<double> nd = node(st, x, y, z);
                    double mr[Q], mb[Q];
                    for (int i = 0; i < Q; ++i) {
                        const int ty = y + kCY[i];
                        mr[i] = mb[i] = 0.0;
                        if (ty < 1 || ty > NY - 2) continue;
                        const size_t t = id(wrap(x + kCX[i], NX), ty, wrap(z + kCZ[i], NZ));
                        mr[i] = lr[size_t(i) * N + t];
                        mb[i] = lb[size_t(i) * N + t];
                    }
                    const AdjOut<double> a = adj_vjp<double>(nd, mr, mb);
                    const size_t g = id(x, y, z);
                    for (int k = 0; k < 11; ++k) LM[size_t(k) * N + g] = a.m[k];
                    for (int k = 0; k < 3; ++k) LF[size_t(k) * N + g] = a.F[k];
                    dp[0] += a.A;
                    dp[1] += a.beta;
                    if (w_uy) LM[3 * N + g] += w_uy[g];
                }
        std::copy(lgB, lgB + 5 * nxz, outgB);
        const double oms_in = 1.0 - 1.0 / c.taub, oms_out = 1.0 - omega;
        for (int z = 1; z < NZ - 1; ++z)
            for (int x = 1; x < NX - 1; ++x) {
                const size_t o = size_t(z) * NX + x;
                // inlet
                if (is_jet(x, z)) {
                    double m[11], e[3], dU[3], da[3], le[3] = {0, 0, 0}, lP[6] = {0, 0, 0, 0, 0, 0}, lu[3] = {0, 0, 0};
                    const size_t g = id(x, 1, z);
                    mom(st, g, m);
                    inlet(n, x, z, e, dU, da);
                    for (int k = 0; k < 9; ++k) {
                        const int i = kUp[k], tx = x + kCX[i], tz = z + kCZ[i];
                        if (!inner(tx, tz)) continue;
                        const double lg = lb[size_t(i) * N + id(tx, 1, tz)];  // f_r = 0 * g
                        double l_rho = 0.0;
                        feq_vjp<double>(i, c.rhob0, e[0], e[1], e[2], lg, l_rho, le);
                        fneqr_vjp<double>(i, m + 5, m[2], m[3], m[4], oms_in * lg, lP, lu);
                    }
                    for (int k = 0; k < 6; ++k) LM[size_t(5 + k) * N + g] += lP[k];
                    for (int k = 0; k < 3; ++k) LM[size_t(2 + k) * N + g] += lu[k];
                    for (int k = 0; k < 3; ++k) {
                        dinlet[size_t(k) * nxz + o] += le[k];
                        dp[2] += le[k] * dU[k];
                        dp[3] += le[k] * da[k];
                    }
                }
                // outlet
                {
                    double m[11], gB1[5], r[9], b[9];
                    const size_t g = id(x, NY - 2, z);
                    mom(st, g, m);
                    outlet_pops(st, gB, x, z, gB1, r, b);
                    const double rhoT = gB1[0] + gB1[1], aR = gB1[0] / rhoT, aB = gB1[1] / rhoT;
                    double l_aR = 0, l_aB = 0, l_rhoT = 0, l_uB[3] = {0, 0, 0}, lP[6] = {0, 0, 0, 0, 0, 0}, lu[3] = {0, 0, 0};
                    for (int k = 0; k < 9; ++k) {
                        const int i = kDown[k], tx = x + kCX[i], tz = z + kCZ[i];
                        if (!inner(tx, tz)) continue;
                        const size_t t = id(tx, NY - 2, tz);
                        const double lrk = lr[size_t(i) * N + t], lbk = lb[size_t(i) * N + t];
                        const double gk = aB > 0 || aR > 0 ? (r[k] + b[k]) : 0.0;  // g = (aR + aB) g
                        l_aR += gk * lrk;
                        l_aB += gk * lbk;
                        const double lg = aR * lrk + aB * lbk;
                        feq_vjp<double>(i, rhoT, gB1[2], gB1[3], gB1[4], lg, l_rhoT, l_uB);
                        fneqr_vjp<double>(i, m + 5, m[2], m[3], m[4], oms_out * lg, lP, lu);
                    }
                    double L[5];
                    const double q = (gB1[0] * l_aR + gB1[1] * l_aB) / (rhoT * rhoT);
                    L[0] = l_aR / rhoT + l_rhoT - q + lgB[o];
                    L[1] = l_aB / rhoT + l_rhoT - q + lgB[nxz + o];
                    for (int k = 0; k < 3; ++k) L[2 + k] = l_uB[k] + lgB[(2 + k) * nxz + o];
                    const double uy = m[3], uc = std::min(std::max(uy, 0.0), 1.0);
                    const double mf[5] = {m[0], m[1], m[2], m[3], m[4]};
                    double l_uc = 0.0;
                    for (int k = 0; k < 5; ++k) {
                        const double gBk = gB[k * nxz + o];
                        outgB[k * nxz + o] = (1.0 - uc) * L[k];
                        LM[size_t(k) * N + g] += uc * L[k];
                        l_uc += -(gBk - mf[k]) * L[k];
                    }
                    if (uy > 0.0 && uy < 1.0) LM[3 * N + g] += l_uc;
                    for (int k = 0; k < 6; ++k) LM[size_t(5 + k) * N + g] += lP[k];
                    for (int k = 0; k < 3; ++k) LM[size_t(2 + k) * N + g] += lu[k];
                }
            }
        // colour gradient transpose: psi of interior nodes and of the outlet ghost plane
        auto Lpsi = [&](int x, int y, int z) {
            double s = 0.0;
            for (int i = 1; i < Q; ++i) {
                const int sx = x - kCX[i], sy = y - kCY[i], sz = z - kCZ[i];
                if (!interior(sx, sy, sz)) continue;
                const size_t q = id(sx, sy, sz);
                s += 3.0 * wT<double>(i) * (kCX[i] * LF[q] + kCY[i] * LF[N + q] + kCZ[i] * LF[2 * N + q]);
            }
            return s;
        };
        for (int z = 1; z < NZ - 1; ++z)
            for (int x = 1; x < NX - 1; ++x) {
                const size_t o = size_t(z) * NX + x;
                const double lp = Lpsi(x, NY - 1, z);
                const double rr = gB[o], rb = gB[nxz + o], rho = rr + rb;
                outgB[o] += lp * 2.0 * rb / (rho * rho);
                outgB[nxz + o] += -lp * 2.0 * rr / (rho * rho);
            }
        std::fill(outr, outr + size_t(Q) * N, 0.0);
        std::fill(outb, outb + size_t(Q) * N, 0.0);
        // stale slots keep their adjoint (the population is carried unchanged)
        for (int z = 1; z < NZ - 1; ++z)
            for (int x = 1; x < NX - 1; ++x)
                for (int i = 0; i < Q; ++i) {
                    const int sx = x - kCX[i], sz = z - kCZ[i];
                    if (kCY[i] == 1 && !(inner(sx, sz) && is_jet(sx, sz))) {
                        const size_t t = id(x, 1, z);
                        outr[size_t(i) * N + t] = lr[size_t(i) * N + t];
                        outb[size_t(i) * N + t] = lb[size_t(i) * N + t];
                    }
                    if (kCY[i] == -1 && !inner(sx, sz)) {
                        const size_t t = id(x, NY - 2, z);
                        outr[size_t(i) * N + t] = lr[size_t(i) * N + t];
                        outb[size_t(i) * N + t] = lb[size_t(i) * N + t];
                    }
                }
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const size_t g = id(x, y, z);
                    double m[11], lam[11];
                    mom(st, g, m);
                    for (int k = 0; k < 11; ++k) lam[k] = LM[size_t(k) * N + g];
                    const double lp = Lpsi(x, y, z), rho = m[0] + m[1];
                    lam[0] += lp * 2.0 * m[1] / (rho * rho);
                    lam[1] += -lp * 2.0 * m[0] / (rho * rho);
                    double r[Q], b[Q];
                    adj_moments<double>(m, lam, r, b);
                    for (int i = 0; i < Q; ++i) {
                        outr[size_t(i) * N + g] += r[i];
                        outb[size_t(i) * N + g] += b[i];
                    }
                }
    }
};

}  // namespace mcg
