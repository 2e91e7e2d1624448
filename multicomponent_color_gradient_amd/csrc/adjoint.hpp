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

// node-layout row of direction i (colour c) at interior node (x,y,z)
__device__ __forceinline__ uint32_t nrow(const DevParams &P, int c, int i, int x, int y, int z) {
    return uint32_t(c * Q + i) * uint32_t(P.PX) + pidx(P, x, y, z);
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

// Adjoint of the collision at every interior node: lam_m (11 planes),
// lam_F (3 planes), and the workgroup's sums of dJ/dA, dJ/dbeta in
// part[2 b], part[2 b + 1] (b = linear workgroup id; reduced by k_adj_sum in
// a fixed order, so the FP64 parameter gradients are deterministic).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MCG_ADJ_WAVES_PER_EU)))
void k_adj_node(DevParams P, const float *__restrict__ mom, const float *__restrict__ ps,
                const float *__restrict__ lam_in, float *__restrict__ lam_m, float *__restrict__ lam_F, double *part) {
    __shared__ double wsum[2][4];
    const int x = 1 + int(blockIdx.x) * 64 + int(threadIdx.x);
    const int y = 1 + int(blockIdx.y) * 4 + int(threadIdx.y);
    const int z = 1 + int(blockIdx.z);
    const bool live = x <= P.NX - 2 && y <= P.NY - 2;
    double gA = 0.0, gB = 0.0;
    if (live) {
        const uint32_t o = midx(P, x, y, z);
        float m[11];
        for (int k = 0; k < 11; ++k) m[k] = mom[size_t(k) * P.plane + o];
        float Fx = 0.f, Fy = 0.f, Fz = 0.f;
        auto phi = [&](auto Ic) -> float {
            constexpr int I = decltype(Ic)::value;
            return ps[midx(P, wrapN(x + kCX[I], P.NX), wrapN(y + kCY[I], P.NY), wrapN(z + kCZ[I], P.NZ))];
        };
        color_gradient(phi, Fx, Fy, Fz);
        const float omega = omega_node(P, y, m[0], m[1]);
        const AdjNode<float> n = adj_setup<float>(m, Fx, Fy, Fz, A_node(P, m[0], m[1]), P.beta, omega);
        // mu_i(x) = lambda_i(x + c_i), loaded where the pair needs it
        const AdjOut<float> a = adj_vjp_fn<float>(n, [&](int i, float &r, float &b) {
            const uint32_t q = pidx(P, wrapN(x + kCX[i], P.NX), wrapN(y + kCY[i], P.NY), wrapN(z + kCZ[i], P.NZ));
            r = lam_in[uint32_t(i) * uint32_t(P.PX) + q];
            b = lam_in[uint32_t(Q + i) * uint32_t(P.PX) + q];
        });
        for (int k = 0; k < 11; ++k) lam_m[size_t(k) * P.plane + o] = a.m[k];
        for (int k = 0; k < 3; ++k) lam_F[size_t(k) * P.plane + o] = a.F[k];
        gA = double(a.A);
        gB = double(a.beta);
    }
    gA = wave_sum(gA);
    gB = wave_sum(gB);
    if (threadIdx.x == 0) {
        wsum[0][threadIdx.y] = gA;
        wsum[1][threadIdx.y] = gB;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const size_t b = size_t(blockIdx.x) + size_t(gridDim.x) * (size_t(blockIdx.y) + size_t(gridDim.y) * blockIdx.z);
        part[2 * b] = (wsum[0][0] + wsum[0][1]) + (wsum[0][2] + wsum[0][3]);
        part[2 * b + 1] = (wsum[1][0] + wsum[1][1]) + (wsum[1][2] + wsum[1][3]);
    }
}

// acc[0..1] += the sums of the nb workgroup partials (one workgroup, fixed order)
__global__ __launch_bounds__(256) void k_adj_sum(const double *__restrict__ part, size_t nb, double *acc) {
    __shared__ double sh[2][256];
    double a = 0.0, b = 0.0;
    for (size_t k = threadIdx.x; k < nb; k += 256) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + h];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[0] += sh[0][0];
        acc[1] += sh[1][0];
    }
}

// Colour-gradient transpose, psi derivative and moments-pass adjoint:
// lambda^n at every interior node.
__global__ __launch_bounds__(256) void k_adj_pop(DevParams P, const float *__restrict__ mom,
                                                 const float *__restrict__ lam_m, const float *__restrict__ lam_F,
                                                 float *__restrict__ lam_out) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const uint32_t o = midx(P, x, y, z);
    float Lpsi = 0.f;
    unroll<1, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        const uint32_t q = midx(P, wrapN(x - kCX[I], P.NX), wrapN(y - kCY[I], P.NY), wrapN(z - kCZ[I], P.NZ));
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
    adj_moments<float>(m, lam, lr, lb);
    for (int i = 0; i < Q; ++i) {
        __builtin_nontemporal_store(lr[i], &lam_out[nrow(P, 0, i, x, y, z)]);
        __builtin_nontemporal_store(lb[i], &lam_out[nrow(P, 1, i, x, y, z)]);
    }
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

    size_t id(int x, int y, int z) const { return size_t(x) + size_t(NX) * (size_t(y) + size_t(NY) * size_t(z)); }
    static int wrap(int v, int n) { return v == 0 ? n - 2 : (v == n - 1 ? 1 : v); }

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
        for (int a = 0; a < 3; ++a) m[2 + a] = j[a] / rho;
        const double *u = m + 2;
        m[5 + PXX] = S[PXX] - rho * u[0] * u[0];
        m[5 + PXY] = S[PXY] - rho * u[0] * u[1];
        m[5 + PYY] = S[PYY] - rho * u[1] * u[1];
        m[5 + PYZ] = S[PYZ] - rho * u[1] * u[2];
        m[5 + PZZ] = S[PZZ] - rho * u[2] * u[2];
        m[5 + PXZ] = S[PXZ] - rho * u[0] * u[2];
    }

    // moments of every interior node (11 arrays of N) and psi
    void all_moments(const double *fr, const double *fb, std::vector<double> &M, std::vector<double> &ps) const {
        M.assign(11 * N, 0.0);
        ps.assign(N, 0.0);
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

    AdjNode<double> node(const std::vector<double> &M, const std::vector<double> &ps, int x, int y, int z) const {
        const size_t g = id(x, y, z);
        double m[11];
        for (int k = 0; k < 11; ++k) m[k] = M[size_t(k) * N + g];
        double F[3] = {0, 0, 0};
        for (int i = 1; i < Q; ++i) {
            const double p = ps[id(wrap(x + kCX[i], NX), wrap(y + kCY[i], NY), wrap(z + kCZ[i], NZ))];
            F[0] += 3.0 * wT<double>(i) * p * kCX[i];
            F[1] += 3.0 * wT<double>(i) * p * kCY[i];
            F[2] += 3.0 * wT<double>(i) * p * kCZ[i];
        }
        return adj_setup<double>(m, F[0], F[1], F[2], A, beta, omega);
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
                        const size_t t = id(wrap(x + kCX[i], NX), wrap(y + kCY[i], NY), wrap(z + kCZ[i], NZ));
                        gr[size_t(i) * N + t] = r[i];
                        gb[size_t(i) * N + t] = b[i];
                    }
                }
    }

    // one adjoint step: lam (adjoint of f^{n+1}) -> out (adjoint of f^n)
    void adjoint(const double *fr, const double *fb, const double *lr, const double *lb, double *outr, double *outb,
                 double &dA, double &dbeta) const {
        std::vector<double> M, ps;
        all_moments(fr, fb, M, ps);
        std::vector<double> LM(11 * N, 0.0), LF(3 * N, 0.0);
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const AdjNode<double> n = node(M, ps, x, y, z);
                    double mr[Q], mb[Q];
                    for (int i = 0; i < Q; ++i) {
                        const size_t t = id(wrap(x + kCX[i], NX), wrap(y + kCY[i], NY), wrap(z + kCZ[i], NZ));
                        mr[i] = lr[size_t(i) * N + t];
                        mb[i] = lb[size_t(i) * N + t];
                    }
                    const AdjOut<double> a = adj_vjp<double>(n, mr, mb);
                    const size_t g = id(x, y, z);
                    for (int k = 0; k < 11; ++k) LM[size_t(k) * N + g] = a.m[k];
                    for (int k = 0; k < 3; ++k) LF[size_t(k) * N + g] = a.F[k];
                    dA += a.A;
                    dbeta += a.beta;
                }
        std::fill(outr, outr + size_t(Q) * N, 0.0);
        std::fill(outb, outb + size_t(Q) * N, 0.0);
        for (int z = 1; z < NZ - 1; ++z)
            for (int y = 1; y < NY - 1; ++y)
                for (int x = 1; x < NX - 1; ++x) {
                    const size_t g = id(x, y, z);
                    double Lpsi = 0;
                    for (int i = 1; i < Q; ++i) {
                        const size_t q = id(wrap(x - kCX[i], NX), wrap(y - kCY[i], NY), wrap(z - kCZ[i], NZ));
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
                    adj_moments<double>(m, lam, r, b);
                    for (int i = 0; i < Q; ++i) {
                        outr[size_t(i) * N + g] = r[i];
                        outb[size_t(i) * N + g] = b[i];
                    }
                }
    }
};

}  // namespace mcg
