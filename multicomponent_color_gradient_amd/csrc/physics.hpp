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

// Accumulators of the moments pass, lbm.cuh:82-165.  The colour densities are
// summed in direction order i = 0..26 (psi of a halo position repeats exactly
// these sums).  The diagonal second moments accumulate g H_aa (H_aa = c_a^2 -
// cs^2) with one FMA each, as the reference does: its terms of both signs keep
// the partial sums near Pi_aa instead of near rho/3, which keeps ~3x more of
// the FP32 precision of Pi^neq than summing g c_a^2 and subtracting cs^2 rho.
// Off-diagonal moments need only additions (H_ab = +-1 or 0).
struct MomentAcc {
    float sr = 0.f, sb = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
    float sxx = 0.f, sxy = 0.f, syy = 0.f, syz = 0.f, szz = 0.f, sxz = 0.f;

    // g: the two-colour population f_r + f_b of direction I (16-bit storage
    // hands over the stored total directly)
    template <int I>
    __device__ __forceinline__ void add(float fr, float fb, float g) {
        sr += fr;
        sb += fb;
        if constexpr (kCX[I] > 0) jx += g; else if constexpr (kCX[I] < 0) jx -= g;
        if constexpr (kCY[I] > 0) jy += g; else if constexpr (kCY[I] < 0) jy -= g;
        if constexpr (kCZ[I] > 0) jz += g; else if constexpr (kCZ[I] < 0) jz -= g;
        sxx = __builtin_fmaf(g, Hxx(I), sxx);
        syy = __builtin_fmaf(g, Hyy(I), syy);
        szz = __builtin_fmaf(g, Hzz(I), szz);
        if constexpr (Hxy(I) > 0.0f) sxy += g; else if constexpr (Hxy(I) < 0.0f) sxy -= g;
        if constexpr (Hyz(I) > 0.0f) syz += g; else if constexpr (Hyz(I) < 0.0f) syz -= g;
        if constexpr (Hxz(I) > 0.0f) sxz += g; else if constexpr (Hxz(I) < 0.0f) sxz -= g;
    }

    // ghalf = gravity/2: with the body force F_y = -gravity rho_r the velocity
    // is (j + F/2)/rho (Guo); ghalf = 0 leaves every value bit-identical
    __device__ __forceinline__ Moments finish(float ghalf) const {
        Moments m;
        m.rr = sr;
        m.rb = sb;
        const float rho = sr + sb;
        const float inv = 1.0f / rho;
        m.ux = jx * inv;
        m.uy = (ghalf != 0.0f ? jy - ghalf * sr : jy) * inv;
        m.uz = jz * inv;
        m.pxx = sxx - rho * m.ux * m.ux;
        m.pxy = sxy - rho * m.ux * m.uy;
        m.pyy = syy - rho * m.uy * m.uy;
        m.pyz = syz - rho * m.uy * m.uz;
        m.pzz = szz - rho * m.uz * m.uz;
        m.pxz = sxz - rho * m.ux * m.uz;
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

// Collision of one node (lbm.cuh:184-291, forces.cuh:50-105,
// collisionOperators.cuh:15-69) in even/odd form.  With c -> -c every term of
// the post-collision population is even (rho, the second-order Hermite terms
// of feq and fneqr, the perturbation Omega2) or odd (the first- and
// third-order terms, the recolouring increment), so the pair (i, opp(i))
// shares one evaluation:
//   g_{+-} = w_i (E +- O),   f_r = aR g + w_i D,   f_b = aB g - w_i D  (D odd)
// with, per node,
//   E = rho (1 - u.u/2cs^2) + rho (c.u)^2/2cs^4 + (1-omega) H2:Pi/2cs^4 + Omega2/w
//   O = rho (c.u)(1/cs^2 + (c.u)^2/6cs^6 - u.u/2cs^4) + (1-omega) H3:T/2cs^6
//   Omega2/w = (A/2)|F| ((c.F)^2/|F|^2 - 1)   (B_i = w_i for i != 0; the rest
//              direction gets (A/2)|F| (10/27)/(8/27))
//   D = beta rr rb/rho^2 rho (c.F)/(|c||F|)
// Bulk nodes (4 rr rb/rho^2 <= 1e-4) carry zero Omega2 and D coefficients, so
// they reduce exactly to f = a g as in the reference's bulk branch, and so do
// the |F| <= 1e-4 cuts of the cos^2 rule and the recolouring.
struct NodeColl {
    float aR, aB;
    float ux, uy, uz;
    float e0, e1, o1, o3;  // feq: E += e0 + e1 (c.u)^2, O += (c.u)(o1 + o3 (c.u)^2)
    float q2, q3;          // (1-omega)/2cs^4, (1-omega)/2cs^6
    float tr;              // -cs^2 tr(Pi): the diagonal remainder of H2:Pi
    float pxx, pyy, pzz, dxy, dyz, dxz;  // Pi diagonal and 2 Pi off-diagonal
    float txxy, txxz, txyy, txzz, tyyz, tyzz, txyz;  // third-order products of fneqr
    float Fx, Fy, Fz;
    float s2, sB;          // Omega2/w = s2 (c.F)^2 - sB
    float rc;              // D = rc (c.F)/|c|
    float g9, g3;          // body force: E += g9 c_y (c.u), O += g3 c_y (g9 = 9 (1 - omega/2) F_y, g3 = 1.5 F_y)
    bool force;            // g9, g3 are used only when set (a constant false drops their code)
};

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// BULK: the node carries no Omega2 or recolouring (s2 = sB = rc = 0, F = 0),
// so those terms, each an exact zero contribution, are left out; the values
// are those of the general form (a zero may differ in sign).
// rest direction: r = aR g0, b = aB g0
template <bool BULK = false>
__device__ __forceinline__ void collide_rest(const NodeColl &n, float &r, float &b) {
    float E = fma_(n.q2, n.tr, n.e0);
    if constexpr (!BULK) E = fma_(n.sB, 1.25f, E);
    const float g = w(0) * E;
    r = n.aR * g;
    b = n.aB * g;
}

// Even and odd parts of the pair (I, I+1) divided by w_I: E, O and the
// recolouring term D (directions 2k-1 and 2k are opposite).
template <int I, bool BULK = false>
__device__ __forceinline__ void pair_terms(const NodeColl &n, float &E, float &O, float &D) {
    static_assert(I % 2 == 1 && kCX[I] == -kCX[I + 1] && kCY[I] == -kCY[I + 1] && kCZ[I] == -kCZ[I + 1], "pair");
    const float cu = cdot<I>(n.ux, n.uy, n.uz);
    const float cu2 = cu * cu;
    float a2 = n.tr;
    if constexpr (kCX[I] != 0) a2 += n.pxx;
    if constexpr (kCY[I] != 0) a2 += n.pyy;
    if constexpr (kCZ[I] != 0) a2 += n.pzz;
    if constexpr (Hxy(I) > 0.0f) a2 += n.dxy; else if constexpr (Hxy(I) < 0.0f) a2 -= n.dxy;
    if constexpr (Hyz(I) > 0.0f) a2 += n.dyz; else if constexpr (Hyz(I) < 0.0f) a2 -= n.dyz;
    if constexpr (Hxz(I) > 0.0f) a2 += n.dxz; else if constexpr (Hxz(I) < 0.0f) a2 -= n.dxz;
    float a3 = 0.0f;
    if constexpr (Hxxy(I) != 0.0f) a3 = fma_(n.txxy, Hxxy(I), a3);
    if constexpr (Hxxz(I) != 0.0f) a3 = fma_(n.txxz, Hxxz(I), a3);
    if constexpr (Hxyy(I) != 0.0f) a3 = fma_(n.txyy, Hxyy(I), a3);
    if constexpr (Hxzz(I) != 0.0f) a3 = fma_(n.txzz, Hxzz(I), a3);
    if constexpr (Hyyz(I) != 0.0f) a3 = fma_(n.tyyz, Hyyz(I), a3);
    if constexpr (Hyzz(I) != 0.0f) a3 = fma_(n.tyzz, Hyzz(I), a3);
    if constexpr (Hxyz(I) != 0.0f) a3 = fma_(n.txyz, Hxyz(I), a3);
    E = fma_(n.e1, cu2, n.e0);
    E = fma_(n.q2, a2, E);
    float d = 0.0f;
    if constexpr (!BULK) {
        d = cdot<I>(n.Fx, n.Fy, n.Fz);
        E = fma_(n.s2, d * d, E) - n.sB;
    }
    O = cu * fma_(n.o3, cu2, n.o1);
    O = fma_(n.q3, a3, O);
    // body force F = (0, F_y, 0): even part (1 - omega/2)(9 (c.u)(c.F) - 3 u.F) (its
    // u-only term sits in e0), odd part 1.5 c.F
    if (n.force) {
        if constexpr (kCY[I] > 0) {
            E = fma_(n.g9, cu, E);
            O = O + n.g3;
        } else if constexpr (kCY[I] < 0) {
            E = fma_(-n.g9, cu, E);
            O = O - n.g3;
        }
    }
    if constexpr (BULK) {
        D = 0.0f;
    } else {
        D = n.rc * d;
        if constexpr (csq(I) > 1) D = D * invc(I);
    }
}

// direction I and its opposite I+1
template <int I, bool BULK = false>
__device__ __forceinline__ void collide_pair(const NodeColl &n, float &rp, float &bp, float &rm, float &bm) {
    float E, O, D;
    pair_terms<I, BULK>(n, E, O, D);
    const float gp = w(I) * (E + O);
    const float gm = w(I) * (E - O);
    if constexpr (BULK) {
        rp = n.aR * gp;
        bp = n.aB * gp;
        rm = n.aR * gm;
        bm = n.aB * gm;
    } else {
        const float Dw = w(I) * D;
        rp = fma_(n.aR, gp, Dw);
        bp = fma_(n.aB, gp, -Dw);
        rm = fma_(n.aR, gm, -Dw);
        bm = fma_(n.aB, gm, Dw);
    }
}

// The same pair for 16-bit storage, which keeps populations divided by w_I:
// the offsets (E +- O) - 1 and the two colours (E +- O) aR +- D, (E +- O) aB -+ D.
template <int I, bool BULK = false>
__device__ __forceinline__ void collide_pair_w(const NodeColl &n, float &ap, float &rp, float &bp, float &am, float &rm,
                                               float &bm) {
    float E, O, D;
    pair_terms<I, BULK>(n, E, O, D);
    const float gp = E + O, gm = E - O;
    const float em1 = E - 1.0f;
    ap = em1 + O;
    am = em1 - O;
    if constexpr (BULK) {
        rp = n.aR * gp;
        bp = n.aB * gp;
        rm = n.aR * gm;
        bm = n.aB * gm;
    } else {
        rp = fma_(n.aR, gp, D);
        bp = fma_(n.aB, gp, -D);
        rm = fma_(n.aR, gm, -D);
        bm = fma_(n.aB, gm, D);
    }
}

// Regularised boundary population (inlet jet/boundary_conditions.cuh:76-79,
// outlet :225-227): g = feq(rho, u_eq) + (1-omega) fneqr(Pi, u_neq).
template <int I>
__device__ __forceinline__ float boundary_pop(float rho, float ex, float ey, float ez, float oms, const Moments &mn) {
    return feq<I>(rho, ex, ey, ez) + oms * fneqr<I>(mn);
}

}  // namespace mcg
