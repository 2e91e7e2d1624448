// kernels.hpp -- two-pass step kernels and the service kernels (init, export,
// statistics, reductions).  The fused step kernel lives in fused.hpp.
#pragma once

#include "device.hpp"

namespace mcg {

// 3-D interior launch: block (64,4,1), grid (PX/64, ceil((NY-2)/4), NZ-2)
#define MCG_INTERIOR_XYZ(P, x, y, z)                 \
    const int x = 1 + int(blockIdx.x) * 64 + int(threadIdx.x); \
    const int y = 1 + int(blockIdx.y) * 4 + int(threadIdx.y);  \
    const int z = 1 + int(blockIdx.z);                          \
    if (x > (P).NX - 2 || y > (P).NY - 2) return;

// ---------------------------------------------------------------- init

struct InitGeo {
    float g[5];          // bubble_x0, bubble_y0, bubble_z0, bubble_radius, interface_width
    const float *sph;    // kind 3: nsph spheres (x, y, z, R)
    int nsph;
};

// initial colour densities, jet/jetIn.cuh:13-62, bubble/bubbleIn.cuh:10-23,
// the tanh bubble and the field of tanh spheres (extensions)
__device__ __forceinline__ void init_rho(const DevParams &P, int kind, const InitGeo &geo, int x, int y, int z,
                                         float &rr, float &rb) {
    if (kind == 0) {
        rr = P.rhor0;
        rb = 0.0f;
        if (P.y_mode == MCG_Y_JET && y == 0 && x >= 1 && x <= P.NX - 2 && z >= 1 && z <= P.NZ - 2) {
            const float j = is_jet(P, x, z) ? 1.0f : 0.0f;
            rr = (1.0f - j) * P.rhor0;
            rb = j * P.rhob0;
        }
        return;
    }
    if (kind == 3) {
        // blue fraction = max over the spheres of (1 + tanh((R_k - r_k)/w))/2
        float frac = 0.0f;
        for (int k = 0; k < geo.nsph; ++k) {
            const float dx = float(x) - geo.sph[4 * k], dy = float(y) - geo.sph[4 * k + 1], dz = float(z) - geo.sph[4 * k + 2];
            const float r = sqrtf(dx * dx + dy * dy + dz * dz);
            frac = fmaxf(frac, (1.0f + tanhf((geo.sph[4 * k + 3] - r) / geo.g[4])) / 2.0f);
        }
        rr = (1.0f - frac) * P.rhor0;
        rb = frac * P.rhob0;
        return;
    }
    const float *g = geo.g;
    const float dx = float(x) - g[0], dy = float(y) - g[1], dz = float(z) - g[2];
    if (kind == 1) {
        const float inb = ((dx * dx) + (dy * dy) + (dz * dz)) <= (g[3] * g[3]) ? 1.0f : 0.0f;
        rr = (1.0f - inb) * P.rhor0;
        rb = inb * P.rhob0;
    } else {
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        const float frac = (1.0f + tanhf((g[3] - r) / g[4])) / 2.0f;
        rr = (1.0f - frac) * P.rhor0;
        rb = frac * P.rhob0;
    }
}

// Populations f_i = w_i rho_k at every interior node, stored at the pull
// source of each slot (initial state of jetDensity / bubble, kernels.cu:13-42).
template <int ST>
__global__ __launch_bounds__(256) void k_init_pull(DevParams P, int kind, InitGeo geo, void *f, float *bin, float *bout) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    float rr, rb;
    init_rho(P, kind, geo, x, y, z, rr, rb);
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        scatter<ST, I>(P, f, bin, bout, x, y, z, w(I) * rr, w(I) * rb);
    });
}

// ghost densities of the full grid and the outlet ghost state
__global__ __launch_bounds__(256) void k_init_ghost(DevParams P, int kind, InitGeo geo, float *grr, float *grb,
                                                    float *gB0, float *gB1) {
    const int x = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int y = int(blockIdx.y) * 4 + int(threadIdx.y);
    const int z = int(blockIdx.z);
    if (x >= P.NX || y >= P.NY) return;
    float rr, rb;
    init_rho(P, kind, geo, x, y, z, rr, rb);
    const size_t g = gidx(P, x, y, z);
    grr[g] = rr;
    grb[g] = rb;
    if (y == P.NY - 1) {
        const size_t nxz = size_t(P.NX) * P.NZ;
        const size_t o = size_t(z) * P.NX + x;
        const float v[5] = {rr, rb, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 5; ++k) {
            gB0[k * nxz + o] = v[k];
            gB1[k * nxz + o] = v[k];
        }
    }
}

// the inlet rewrites the ghost densities of plane y=0 every step
// (jet/boundary_conditions.cuh:35-36); they only change on the first step
__global__ __launch_bounds__(256) void k_inlet_ghost_reset(DevParams P, float *grr, float *grb) {
    const int x = 1 + int(blockIdx.x) * 256 + int(threadIdx.x);
    const int z = 1 + int(blockIdx.y);
    if (x > P.NX - 2) return;
    const float j = is_jet(P, x, z) ? 1.0f : 0.0f;
    const size_t g = gidx(P, x, 0, z);
    grr[g] = (1.0f - j) * P.rhor0;
    grb[g] = j * P.rhob0;
}

// ---------------------------------------------------------------- two-pass step

// Pass 1: gather + moments (Mfields, kernels.cu:60-76, lbm.cuh:82-165)
template <int ST>
__global__ __launch_bounds__(256) void k_moments(DevParams P, const void *__restrict__ f, const float *__restrict__ bin,
                                                 const float *__restrict__ bout, float *__restrict__ mom) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const Moments m = gather_moments<ST>(P, f, bin, bout, x, y, z);
    const uint32_t p = midx(P, x, y, z);
    const size_t pl = P.plane;
    mom[0 * pl + p] = m.rr;
    mom[1 * pl + p] = m.rb;
    mom[2 * pl + p] = m.ux;
    mom[3 * pl + p] = m.uy;
    mom[4 * pl + p] = m.uz;
    mom[5 * pl + p] = m.pxx;
    mom[6 * pl + p] = m.pxy;
    mom[7 * pl + p] = m.pyy;
    mom[8 * pl + p] = m.pyz;
    mom[9 * pl + p] = m.pzz;
    mom[10 * pl + p] = m.pxz;
}

__device__ __forceinline__ Moments load_moments(const DevParams &P, const float *__restrict__ mom, uint32_t p) {
    const size_t pl = P.plane;
    Moments m;
    m.rr = mom[0 * pl + p];
    m.rb = mom[1 * pl + p];
    m.ux = mom[2 * pl + p];
    m.uy = mom[3 * pl + p];
    m.uz = mom[4 * pl + p];
    m.pxx = mom[5 * pl + p];
    m.pxy = mom[6 * pl + p];
    m.pyy = mom[7 * pl + p];
    m.pyz = mom[8 * pl + p];
    m.pzz = mom[9 * pl + p];
    m.pxz = mom[10 * pl + p];
    return m;
}

// Pass 2: collision, perturbation, recolouring; post-collision populations
// stored at the node (ColliStream, kernels.cu:78-94, lbm.cuh:184-293), plus
// the inlet (inlet, kernels.cu:98-114) and outlet (neumann, kernels.cu:116-132)
// populations of the columns this node feeds.
template <int ST>
__global__ __launch_bounds__(256) void k_collide(DevParams P, DevBufs B) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const uint32_t p = pidx(P, x, y, z);
    const Moments m = load_moments(P, B.mom, midx(P, x, y, z));
    auto phi = [&](auto Ic) -> float {
        constexpr int I = decltype(Ic)::value;
        int xn, yn, zn;
        grad_neighbor(P, x + kCX[I], y + kCY[I], z + kCZ[I], xn, yn, zn);
        if (interior(P, xn, yn, zn)) {
            const uint32_t q = midx(P, xn, yn, zn);
            return psi(B.mom[q], B.mom[size_t(P.plane) + q]);
        }
        return ghost_psi(P, B, xn, yn, zn);
    };
    const NodeColl n = node_setup(P, m, y, phi);
    collide_store<ST>(n, plane_ptr(P, B.f_dst, z), uint32_t(P.PX), pinplane(P, x, y), p, sr_step_key(B.step));
    if (P.y_mode == MCG_Y_JET) {
        if (y == 1 && is_jet(P, x, z)) inlet_emit(P, B.in_dst, m, x, z, B.step, B.inlet_du);
        if (y == P.NY - 2) outlet_emit(P, B.gB_src, B.gB_dst, B.out_dst, m, x, z);
    }
}

// ---------------------------------------------------------------- statistics

// Radial moments of u_y, u_r (and psi) per (y, r-bin), kernels.cu:188-233.
// One block per (y, chunk of z): LDS privatised fixed-point bins (stat_sample),
// then one global integer atomic per touched bin and moment into this step's
// sums; k_stats_fold adds them to the FP64 sums.
constexpr int kRadialZChunk = 8;

__device__ __forceinline__ void radial_flush(const DevParams &P, int y, const unsigned long long *lds,
                                             unsigned long long *stats_fx, unsigned long long *count) {
    const size_t nb = size_t(P.NY) * P.NR;
    for (int b = threadIdx.x; b < P.NR; b += blockDim.x) {
        const unsigned long long c = lds[7 * P.NR + b];
        if (c == 0) continue;
        const size_t o = size_t(y) * P.NR + b;
        for (int k = 0; k < 7; ++k) atomicAdd(&stats_fx[k * nb + o], lds[k * P.NR + b]);
        atomicAdd(&count[o], c);
    }
}

__device__ __forceinline__ void radial_add(const DevParams &P, int x, int z, const Moments &m, unsigned long long *lds) {
    float dx, dz, r;
    const int b = radial_bin(P, x, z, dx, dz, r);
    if (b < 0) return;
    float ur = 0.0f;
    if (r > 1.0e-12f) ur = (m.ux * dx + m.uz * dz) / r;
    stat_sample(lds + b, size_t(P.NR), double(m.uy), double(ur), double(psi(m.rr, m.rb)));
}

__global__ __launch_bounds__(256) void k_radial(DevParams P, const float *__restrict__ mom,
                                                unsigned long long *stats_fx, unsigned long long *count) {
    extern __shared__ unsigned long long lds[];  // [8][NR]: 7 fixed-point sums, then the counts
    for (int b = threadIdx.x; b < 8 * P.NR; b += blockDim.x) lds[b] = 0ull;
    __syncthreads();
    const int y = 1 + int(blockIdx.x);
    const int z0 = 1 + int(blockIdx.y) * kRadialZChunk;
    const int ix = P.NX - 2;
    const int nz = min(kRadialZChunk, P.NZ - 1 - z0);
    for (int t = threadIdx.x; t < ix * nz; t += blockDim.x) {
        const int x = 1 + t % ix;
        const int z = z0 + t / ix;
        const uint32_t p = midx(P, x, y, z);
        Moments m;
        m.rr = mom[p];
        m.rb = mom[size_t(P.plane) + p];
        m.ux = mom[2 * size_t(P.plane) + p];
        m.uy = mom[3 * size_t(P.plane) + p];
        m.uz = mom[4 * size_t(P.plane) + p];
        radial_add(P, x, z, m, lds);
    }
    __syncthreads();
    radial_flush(P, y, lds, stats_fx, count);
}

// Adds the fixed-point sums of one statistics step to the FP64 running sums and
// clears them for the next step: one thread per (y, bin), so each FP64 sum
// takes one add per step in step order and the sums are reproducible.
__global__ __launch_bounds__(256) void k_stats_fold(size_t nb, unsigned long long *__restrict__ stats_fx,
                                                    double *__restrict__ stats) {
    const size_t o = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (o >= nb) return;
    for (int q = 0; q < 7; ++q) {
        const long long a = (long long)stats_fx[q * nb + o];
        if (a != 0) {
            stats[q * nb + o] += double(a) * (1.0 / kStatScale);
            stats_fx[q * nb + o] = 0ull;
        }
    }
}

// ---------------------------------------------------------------- diagnostics

// Per-block FP64 partial sums over interior nodes from the moment planes:
// [0] sum 0.5|u|^2 (compute_total_tke, kernels.cu:136-152), [1] sum rho_r,
// [2] sum rho_b, [3..5] sum rho u, [6] number of unhealthy nodes.
constexpr int kRedQ = 7;

__global__ __launch_bounds__(256) void k_reduce_partials(DevParams P, const float *__restrict__ mom, double *partials,
                                                         int nodes_per_block) {
    __shared__ double sh[kRedQ][256];
    double acc[kRedQ] = {0, 0, 0, 0, 0, 0, 0};
    const size_t total = size_t(P.NX - 2) * (P.NY - 2) * (P.NZ - 2);
    const size_t start = size_t(blockIdx.x) * nodes_per_block;
    const size_t end = min(total, start + size_t(nodes_per_block));
    const size_t ix = P.NX - 2, iy = P.NY - 2;
    for (size_t t = start + threadIdx.x; t < end; t += blockDim.x) {
        const int x = 1 + int(t % ix);
        const int y = 1 + int((t / ix) % iy);
        const int z = 1 + int(t / (ix * iy));
        const uint32_t p = midx(P, x, y, z);
        const float rr = mom[p], rb = mom[size_t(P.plane) + p];
        const float vx = mom[2 * size_t(P.plane) + p], vy = mom[3 * size_t(P.plane) + p], vz = mom[4 * size_t(P.plane) + p];
        const float ke = 0.5f * (vx * vx + vy * vy + vz * vz);  // meanF/meanFields.cuh:10-19
        acc[0] += double(ke);
        acc[1] += double(rr);
        acc[2] += double(rb);
        const double rho = double(rr) + double(rb);
        acc[3] += rho * double(vx);
        acc[4] += rho * double(vy);
        acc[5] += rho * double(vz);
        const bool bad = !(rr >= 0.0f) || !(rb >= 0.0f) || !(rr + rb > 0.0f) || !isfinite(rr + rb) ||
                         !isfinite(vx) || !isfinite(vy) || !isfinite(vz);
        acc[6] += bad ? 1.0 : 0.0;
    }
    for (int k = 0; k < kRedQ; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < kRedQ; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < kRedQ; ++k) partials[size_t(blockIdx.x) * kRedQ + k] = sh[k][0];
}

// fixed-order final reduction (deterministic)
__global__ __launch_bounds__(256) void k_reduce_final(const double *partials, int nblocks, double *out) {
    __shared__ double sh[kRedQ][256];
    double acc[kRedQ] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int k = 0; k < kRedQ; ++k) acc[k] += partials[size_t(b) * kRedQ + k];
    for (int k = 0; k < kRedQ; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < kRedQ; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < kRedQ; ++k) out[k] = sh[k][0];
}

// ---------------------------------------------------------------- exchange

// rho_r, rho_b, u of the full grid (VTI layout, io/save_data.cuh:137-200)
__global__ __launch_bounds__(256) void k_export_fields(DevParams P, const float *__restrict__ mom, int zero_interior,
                                                       const float *__restrict__ gB, const float *__restrict__ grr,
                                                       const float *__restrict__ grb, float *out) {
    const int x = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int y = int(blockIdx.y) * 4 + int(threadIdx.y);
    const int z = int(blockIdx.z);
    if (x >= P.NX || y >= P.NY) return;
    const size_t N = size_t(P.NX) * P.NY * P.NZ;
    const size_t g = gidx(P, x, y, z);
    float v[5];
    if (interior(P, x, y, z)) {
        const uint32_t p = midx(P, x, y, z);
        for (int k = 0; k < 5; ++k) v[k] = zero_interior ? 0.0f : mom[k * size_t(P.plane) + p];
    } else if (P.y_mode == MCG_Y_JET && y == P.NY - 1 && x >= 1 && x <= P.NX - 2 && z >= 1 && z <= P.NZ - 2) {
        const size_t nxz = size_t(P.NX) * P.NZ;
        const size_t o = size_t(z) * P.NX + x;
        for (int k = 0; k < 5; ++k) v[k] = gB[k * nxz + o];
    } else {
        v[0] = grr[g];
        v[1] = grb[g];
        v[2] = v[3] = v[4] = 0.0f;
    }
    for (int k = 0; k < 5; ++k) out[k * N + g] = v[k];
}

// populations the next step reads, reference layout i*Ncells + id
template <int ST>
__global__ __launch_bounds__(256) void k_gather_pops(DevParams P, const void *__restrict__ f, const float *__restrict__ bin,
                                                     const float *__restrict__ bout, float *out_r, float *out_b) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const size_t N = size_t(P.NX) * P.NY * P.NZ;
    const size_t g = gidx(P, x, y, z);
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b;
        gather<ST, I>(P, f, bin, bout, x, y, z, a, b);
        out_r[size_t(I) * N + g] = a;
        out_b[size_t(I) * N + g] = b;
    });
}

template <int ST>
__global__ __launch_bounds__(256) void k_scatter_pops(DevParams P, void *f, float *bin, float *bout,
                                                      const float *__restrict__ in_r, const float *__restrict__ in_b) {
    MCG_INTERIOR_XYZ(P, x, y, z)
    const size_t N = size_t(P.NX) * P.NY * P.NZ;
    const size_t g = gidx(P, x, y, z);
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        scatter<ST, I>(P, f, bin, bout, x, y, z, in_r[size_t(I) * N + g], in_b[size_t(I) * N + g]);
    });
}

}  // namespace mcg
