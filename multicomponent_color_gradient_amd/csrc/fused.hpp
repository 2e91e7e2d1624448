// fused.hpp -- one-launch step kernel (placeholder generation: per-node
// gather, neighbour densities re-gathered for interface nodes).
#pragma once

#include "kernels.hpp"

namespace mcg {

// density-based psi of interior neighbour (x,y,z), summed in the order of
// lbm.cuh:106-115 so that it equals psi of the moments pass bit for bit
__device__ __forceinline__ float gathered_psi(const DevParams &P, const DevBufs &B, int x, int y, int z) {
    float sr = 0.f, sb = 0.f;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b;
        gather<I>(P, B.f_src, B.in_src, B.out_src, x, y, z, a, b);
        sr += a;
        sb += b;
    });
    return psi(sr, sb);
}

__global__ __launch_bounds__(256) void k_fused_simple(DevParams P, DevBufs B) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const Moments m = gather_moments(P, B.f_src, B.in_src, B.out_src, x, y, z);
    auto phi = [&](auto Ic) -> float {
        constexpr int I = decltype(Ic)::value;
        int xn, yn, zn;
        grad_neighbor(P, x + kCX[I], y + kCY[I], z + kCZ[I], xn, yn, zn);
        if (interior(P, xn, yn, zn)) return gathered_psi(P, B, xn, yn, zn);
        return ghost_psi(P, B, xn, yn, zn);
    };
    const NodeColl n = node_setup(P, m, y, phi);
    const uint32_t p = pidx(P, x, y, z);
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b;
        collide_dir<I>(n, a, b);
        B.f_dst[size_t(I) * P.plane + p] = a;
        B.f_dst[size_t(Q + I) * P.plane + p] = b;
    });
    if (P.y_mode == MCG_Y_JET) {
        if (y == 1 && is_jet(P, x, z)) inlet_emit(P, B.in_dst, m, x, z);
        if (y == P.NY - 2) outlet_emit(P, B.gB_src, B.gB_dst, B.out_dst, m, x, z);
    }
    if (B.do_stats) {
        float dx, dz, r;
        const int rbin = radial_bin(P, x, z, dx, dz, r);
        if (rbin >= 0) {
            float ur = 0.0f;
            if (r > 1.0e-12f) ur = (m.ux * dx + m.uz * dz) / r;
            const double vy = double(m.uy), vr = double(ur), ph = double(psi(m.rr, m.rb));
            const size_t nb = size_t(P.NY) * P.NR;
            const size_t o = size_t(y) * P.NR + rbin;
            atomicAdd(&B.stats[0 * nb + o], vy);
            atomicAdd(&B.stats[1 * nb + o], vy * vy);
            atomicAdd(&B.stats[2 * nb + o], vr);
            atomicAdd(&B.stats[3 * nb + o], vr * vr);
            atomicAdd(&B.stats[4 * nb + o], vr * vy);
            atomicAdd(&B.stats[5 * nb + o], ph);
            atomicAdd(&B.stats[6 * nb + o], ph * ph);
            atomicAdd(&B.count[o], 1ull);
        }
    }
}

// the placeholder kernel leaves the statistics to k_radial
inline bool fused_has_stats() { return false; }

inline int launch_fused(const DevParams &P, const DevBufs &B, hipStream_t stream) {
    const dim3 grid(unsigned(P.PX / 64), unsigned((P.IY + 3) / 4), unsigned(P.IZ));
    hipLaunchKernelGGL(k_fused_simple, grid, dim3(64, 4, 1), 0, stream, P, B);
    return int(hipGetLastError());
}

}  // namespace mcg
