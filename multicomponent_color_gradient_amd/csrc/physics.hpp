// physics.hpp -- per-node operators of the colour-gradient step, FP32.
//
// Each operator keeps the arithmetic of the reference function it restates
// (cited per function) so that results agree to FP32 rounding; terms whose
// compile-time coefficient is exactly 0 or +-1 are dropped or folded, which
// leaves every finite result unchanged.  The direction index is a template
// parameter so that all lattice coefficients are immediates.
#pragma once

#include <hip/hip_runtime.h>

#include "lattice.hpp"

namespace mcg {

struct Moments {
    float rr, rb;           // colour densities
    float ux, uy, uz;       // barycentric velocity
    float pxx, pxy, pyy, pyz, pzz, pxz;  // non-equilibrium second moment
};

// sum_a v_a c_{i,a} with the zero components dropped
template <int I>
__device__ __forceinline__ float cdot(float x, float y, float z) {
    float s = 0.0f;
    bool any = false;
    if constexpr (kCX[I] != 0) { s = kCX[I] > 0 ? x : -x; any = true; }
    if constexpr (kCY[I] != 0) { const float t = kCY[I] > 0 ? y : -y; s = any ? s + t : t; any = true; }
    if constexpr (kCZ[I] != 0) { const float t = kCZ[I] > 0 ? z : -z; s = any ? s + t : t; any = true; }
    return s;
}

// Third-order Hermite equilibrium, lbm.cuh:15-35.
template <int I>
__device__ __forceinline__ float feq(float rho, float ux, float uy, float uz) {
    const float cu = cdot<I>(ux, uy, uz);
    const float usq = ux * ux + uy * uy + uz * uz;
    const float cu2 = cu * cu;
    const float A2 = (cu * cu) * inv_2cs4 - usq * inv_2cs2 + cu * inv_cs2;
    const float A3 = cu * (cu2 * inv_6cs6 - usq * inv_2cs4);
    return w(I) * rho * (1.0f + A2 + A3);
}

// Regularised non-equilibrium population from Pi^neq and u, lbm.cuh:37-78.
template <int I>
__device__ __forceinline__ float fneqr(const Moments &m) {
    // second order: diagonal Hermite terms are never zero, off-diagonal ones are 0 or +-1
    float A2 = m.pxx * Hxx(I);
    if constexpr (Hxy(I) != 0.0f) A2 += (2.0f * m.pxy) * Hxy(I);
    A2 += m.pyy * Hyy(I);
    if constexpr (Hyz(I) != 0.0f) A2 += (2.0f * m.pyz) * Hyz(I);
    A2 += m.pzz * Hzz(I);
    if constexpr (Hxz(I) != 0.0f) A2 += (2.0f * m.pxz) * Hxz(I);
    A2 *= inv_2cs4;

    float A3 = 0.0f;
    bool any = false;
    auto acc = [&](float a, float h) {
        A3 = any ? A3 + a * h : a * h;
        any = true;
    };
    if constexpr (Hxxy(I) != 0.0f) acc(m.pxx * m.uy + 2.0f * m.pxy * m.ux, Hxxy(I));
    if constexpr (Hxxz(I) != 0.0f) acc(m.pxx * m.uz + 2.0f * m.pxz * m.ux, Hxxz(I));
    if constexpr (Hxyy(I) != 0.0f) acc(m.pyy * m.ux + 2.0f * m.pxy * m.uy, Hxyy(I));
    if constexpr (Hxzz(I) != 0.0f) acc(m.pzz * m.ux + 2.0f * m.pxz * m.uz, Hxzz(I));
    if constexpr (Hyyz(I) != 0.0f) acc(m.pyy * m.uz + 2.0f * m.pyz * m.uy, Hyyz(I));
    if constexpr (Hyzz(I) != 0.0f) acc(m.pzz * m.uy + 2.0f * m.pyz * m.uz, Hyzz(I));
    if constexpr (Hxyz(I) != 0.0f) acc(2.0f * (m.pxy * m.uz + m.pxz * m.uy + m.pyz * m.ux), Hxyz(I));
    A3 *= inv_2cs6;
    return w(I) * (A2 + A3);
}

// Accumulators of the moments pass, lbm.cuh:82-165 (sum order i = 0..26).
struct MomentAcc {
    float sr = 0.f, sb = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
    float axx = 0.f, axy = 0.f, ayy = 0.f, ayz = 0.f, azz = 0.f, axz = 0.f;

    template <int I>
    __device__ __forceinline__ void add(float fr, float fb) {
        sr += fr;
        sb += fb;
        const float g = fr + fb;
        if constexpr (kCX[I] > 0) jx += g; else if constexpr (kCX[I] < 0) jx -= g;
        if constexpr (kCY[I] > 0) jy += g; else if constexpr (kCY[I] < 0) jy -= g;
        if constexpr (kCZ[I] > 0) jz += g; else if constexpr (kCZ[I] < 0) jz -= g;
        axx += g * Hxx(I);
        if constexpr (Hxy(I) > 0.0f) axy += g; else if constexpr (Hxy(I) < 0.0f) axy -= g;
        ayy += g * Hyy(I);
        if constexpr (Hyz(I) > 0.0f) ayz += g; else if constexpr (Hyz(I) < 0.0f) ayz -= g;
        azz += g * Hzz(I);
        if constexpr (Hxz(I) > 0.0f) axz += g; else if constexpr (Hxz(I) < 0.0f) axz -= g;
    }

    __device__ __forceinline__ Moments finish() const {
        Moments m;
        m.rr = sr;
        m.rb = sb;
        const float rho = sr + sb;
        const float inv = 1.0f / rho;
        m.ux = jx * inv;
        m.uy = jy * inv;
        m.uz = jz * inv;
        m.pxx = axx - rho * m.ux * m.ux;
        m.pxy = axy - rho * m.ux * m.uy;
        m.pyy = ayy - rho * m.uy * m.uy;
        m.pyz = ayz - rho * m.uy * m.uz;
        m.pzz = azz - rho * m.uz * m.uz;
        m.pxz = axz - rho * m.ux * m.uz;
        return m;
    }
};

// phase field psi, utilities/mathUtilities.cuh:24-28
__device__ __forceinline__ float psi(float self, float other) { return (self - other) / (self + other); }

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// Colour gradient F = (1/cs^2) sum_{i>=1} w_i psi(x+c_i) c_i, forces.cuh:13-43.
// phi(I) returns psi at the neighbour in direction I.
template <class PhiFn>
__device__ __forceinline__ void color_gradient(PhiFn &&phi, float &Fx, float &Fy, float &Fz) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    unroll<1, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        const float wp = w(I) * phi(Ic);
        if constexpr (kCX[I] > 0) sx += wp; else if constexpr (kCX[I] < 0) sx -= wp;
        if constexpr (kCY[I] > 0) sy += wp; else if constexpr (kCY[I] < 0) sy -= wp;
        if constexpr (kCZ[I] > 0) sz += wp; else if constexpr (kCZ[I] < 0) sz -= wp;
    });
    Fx = inv_cs2 * sx;
    Fy = inv_cs2 * sy;
    Fz = inv_cs2 * sz;
}

// Perturbation operator with the cos^2 rule, collisionOperators.cuh:15-24 and
// forces.cuh:50-75.
template <int I>
__device__ __forceinline__ float omega2(float Fx, float Fy, float Fz, float absF, float A) {
    float c2 = 0.0f;
    if constexpr (I != 0) {
        if (!(absF <= 1.0e-4f)) {
            const float d = cdot<I>(Fx, Fy, Fz);
            c2 = (d * d) / (absF * absF);
        }
    }
    return (A / 2.0f) * absF * (w(I) * c2 - Bc(I));
}

// Latva-Kokko recolouring increment, collisionOperators.cuh:28-69.
// coeff = beta * rr*rb/rho^2 and invF = 1/|F| are per node.
template <int I>
__device__ __forceinline__ float recolor(float rho, float coeff, float Fx, float Fy, float Fz, float absF, float invF) {
    if constexpr (I == 0) {
        return 0.0f;
    } else {
        float c = 0.0f;
        if (!(absF <= 1.0e-4f)) c = cdot<I>(Fx, Fy, Fz) * invF * invc(I);
        const float fe = w(I) * rho;
        return coeff * c * fe;
    }
}

// Node-level collision inputs shared by all 27 directions (lbm.cuh:184-250).
struct NodeColl {
    Moments m;
    float rT, aR, aB, oms;
    bool bulk;
    float Fx, Fy, Fz, absF, invF, A, coeff;
};

// Post-collision populations of direction I, lbm.cuh:218-245 (bulk) and
// :260-291 (interface): g = feq + (1-omega) fneqr [+ Omega2], f_r = aR g [+ D],
// f_b = aB g [- D].
template <int I>
__device__ __forceinline__ void collide_dir(const NodeColl &n, float &outr, float &outb) {
    float g = feq<I>(n.rT, n.m.ux, n.m.uy, n.m.uz) + n.oms * fneqr<I>(n.m);
    if (n.bulk) {
        outr = n.aR * g;
        outb = n.aB * g;
    } else {
        g = g + omega2<I>(n.Fx, n.Fy, n.Fz, n.absF, n.A);
        const float d = recolor<I>(n.rT, n.coeff, n.Fx, n.Fy, n.Fz, n.absF, n.invF);
        outr = n.aR * g + d;
        outb = n.aB * g - d;
    }
}

// Regularised boundary population (inlet jet/boundary_conditions.cuh:128-147,
// outlet :278-296): g = feq(rho, u_eq) + (1-omega) fneqr(Pi, u_neq).
template <int I>
__device__ __forceinline__ float boundary_pop(float rho, float ex, float ey, float ez, float oms, const Moments &mn) {
    return feq<I>(rho, ex, ey, ez) + oms * fneqr<I>(mn);
}

}  // namespace mcg
