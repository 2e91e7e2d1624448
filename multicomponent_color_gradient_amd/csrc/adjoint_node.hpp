// adjoint_node.hpp -- per-node forward and vector-Jacobian product of the
// colour-gradient step, templated on the arithmetic type so that the same
// code serves the FP32 GPU adjoint and the FP64 CPU adjoint.
//
// The forward map of one node y (lbm.cuh:184-291, forces.cuh:13-105,
// collisionOperators.cuh:15-69) in even/odd form (see physics.hpp):
//   inputs : moments m(y) = (rr, rb, u, Pi^neq) and the colour gradient F(y)
//   outputs: the 54 post-collision populations
//   rest    g0 = w0 (rho(1 - 1.5 u.u) + q2 tr(Pi)(-1/3) + 1.25 sB)
//   pair c  E = rho(1 - 1.5 u.u) + 4.5 rho (c.u)^2 + q2 H2:Pi + s2 (c.F)^2 - sB
//           O = rho (c.u)(3 - 4.5 u.u) + 4.5 rho (c.u)^3 + q3 H3:T(Pi,u)
//           D = rc (c.F)/|c|
//           f_r(+-c) = aR w (E +- O) +- w D,  f_b(+-c) = aB w (E +- O) -+ w D
// with q2 = 4.5 (1-omega), q3 = 13.5 (1-omega), sB = (A/2)|F|,
// s2 = (A/2)/|F|, rc = beta rr rb/(rho |F|) (s2 = rc = 0 when |F| <= 1e-4;
// sB = s2 = rc = 0 and F unused at bulk nodes, 4 rr rb/rho^2 <= 1e-4).
// The adjoint treats the bulk switch and the |F| cut as fixed at the forward
// state (the map is differentiated away from its clamps).
// Extensions (AdjExt): the per-node relaxation rate omega = 1/tau(psi),
// tau = (1+psi)/2 taur + (1-psi)/2 taub (q2, q3 then depend on the node's
// densities), and the body force F_y = -grav rr on red: e0 gains
// -3 (1 - omega/2) F_y u_y, E gains 9 (1 - omega/2) F_y c_y (c.u), O gains
// 1.5 F_y c_y; the moments pass shifts u_y by -grav rr/(2 rho).
#pragma once

#include "lattice.hpp"

#if defined(__HIP__)
#define MCG_HD __host__ __device__ inline
#else
#define MCG_HD inline
#endif

namespace mcg {

// indices of the six Pi components and seven third-order products
enum { PXX = 0, PXY, PYY, PYZ, PZZ, PXZ };
enum { TXXY = 0, TXXZ, TXYY, TXZZ, TYYZ, TYZZ, TXYZ };

// Lattice constants evaluated in the arithmetic type: for float they equal
// the FP32 tables of lattice.hpp (what the GPU step uses), for double they are
// the FP64 values of the oracle.
template <class T>
MCG_HD T cs2T() {
    return T(1) / T(3);
}
template <class T>
MCG_HD T wT(int i) {
    const int s = csq(i);
    return s == 0 ? T(8) / T(27) : s == 1 ? T(2) / T(27) : s == 2 ? T(1) / T(54) : T(1) / T(216);
}
template <class T>
MCG_HD T invcT(int i) {
    const int s = csq(i);
    return s == 0 ? T(0) : T(1) / sqrt(T(s));
}
// second-order Hermite coefficient of Pi component k for direction i
template <class T>
MCG_HD T h2T(int k, int i) {
    const T cx = T(kCX[i]), cy = T(kCY[i]), cz = T(kCZ[i]), c = cs2T<T>();
    switch (k) {
        case PXX: return cx * cx - c;
        case PXY: return cx * cy;
        case PYY: return cy * cy - c;
        case PYZ: return cy * cz;
        case PZZ: return cz * cz - c;
        default: return cx * cz;
    }
}
// third-order Hermite coefficient of product k for direction i
template <class T>
MCG_HD T h3T(int k, int i) {
    const T cx = T(kCX[i]), cy = T(kCY[i]), cz = T(kCZ[i]), c = cs2T<T>();
    switch (k) {
        case TXXY: return cx * cx * cy - c * cy;
        case TXXZ: return cx * cx * cz - c * cz;
        case TXYY: return cx * cy * cy - c * cx;
        case TXZZ: return cx * cz * cz - c * cx;
        case TYYZ: return cy * cy * cz - c * cz;
        case TYZZ: return cy * cz * cz - c * cy;
        default: return cx * cy * cz;
    }
}

// Adjoint of the third-order products t_k(Pi, u) of fneqr (lbm.cuh:58-72):
// adds into l_P and l_u.
template <class T>
MCG_HD void t_vjp(const T P[6], T ux, T uy, T uz, const T l_t[7], T l_P[6], T l_u[3]) {
    l_P[PXX] += l_t[TXXY] * uy + l_t[TXXZ] * uz;
    l_P[PYY] += l_t[TXYY] * ux + l_t[TYYZ] * uz;
    l_P[PZZ] += l_t[TXZZ] * ux + l_t[TYZZ] * uy;
    l_P[PXY] += T(2) * (l_t[TXXY] * ux + l_t[TXYY] * uy + l_t[TXYZ] * uz);
    l_P[PXZ] += T(2) * (l_t[TXXZ] * ux + l_t[TXZZ] * uz + l_t[TXYZ] * uy);
    l_P[PYZ] += T(2) * (l_t[TYYZ] * uy + l_t[TYZZ] * uz + l_t[TXYZ] * ux);
    l_u[0] += l_t[TXXY] * T(2) * P[PXY] + l_t[TXXZ] * T(2) * P[PXZ] + l_t[TXYY] * P[PYY] + l_t[TXZZ] * P[PZZ] +
              l_t[TXYZ] * T(2) * P[PYZ];
    l_u[1] += l_t[TXXY] * P[PXX] + l_t[TXYY] * T(2) * P[PXY] + l_t[TYYZ] * T(2) * P[PYZ] + l_t[TYZZ] * P[PZZ] +
              l_t[TXYZ] * T(2) * P[PXZ];
    l_u[2] += l_t[TXXZ] * P[PXX] + l_t[TXZZ] * T(2) * P[PXZ] + l_t[TYYZ] * P[PYY] + l_t[TYZZ] * T(2) * P[PYZ] +
              l_t[TXYZ] * T(2) * P[PXY];
}

// Equilibrium of direction i in the reference form (lbm.cuh:15-35) and its
// derivatives: dfeq/drho and dfeq/du_a.
template <class T>
MCG_HD T feqT(int i, T rho, T ux, T uy, T uz) {
    const T cu = T(kCX[i]) * ux + T(kCY[i]) * uy + T(kCZ[i]) * uz;
    const T usq = ux * ux + uy * uy + uz * uz;
    return wT<T>(i) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq + cu * (T(4.5) * cu * cu - T(4.5) * usq));
}
template <class T>
MCG_HD void feq_vjp(int i, T rho, T ux, T uy, T uz, T lam, T &l_rho, T l_u[3]) {
    const T cu = T(kCX[i]) * ux + T(kCY[i]) * uy + T(kCZ[i]) * uz;
    const T usq = ux * ux + uy * uy + uz * uz;
    const T w = wT<T>(i);
    l_rho += lam * w * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq + cu * (T(4.5) * cu * cu - T(4.5) * usq));
    const T ac = w * rho * (T(3) + T(9) * cu + T(13.5) * cu * cu - T(4.5) * usq);  // coefficient of c_a
    const T au = w * rho * (-T(3) - T(9) * cu);                                  // coefficient of u_a
    l_u[0] += lam * (ac * T(kCX[i]) + au * ux);
    l_u[1] += lam * (ac * T(kCY[i]) + au * uy);
    l_u[2] += lam * (ac * T(kCZ[i]) + au * uz);
}

// Regularised non-equilibrium population (lbm.cuh:37-78) and its adjoint.
template <class T>
MCG_HD void fneqr_products(const T P[6], T ux, T uy, T uz, T t[7]) {
    t[TXXY] = P[PXX] * uy + T(2) * P[PXY] * ux;
    t[TXXZ] = P[PXX] * uz + T(2) * P[PXZ] * ux;
    t[TXYY] = P[PYY] * ux + T(2) * P[PXY] * uy;
    t[TXZZ] = P[PZZ] * ux + T(2) * P[PXZ] * uz;
    t[TYYZ] = P[PYY] * uz + T(2) * P[PYZ] * uy;
    t[TYZZ] = P[PZZ] * uy + T(2) * P[PYZ] * uz;
    t[TXYZ] = T(2) * (P[PXY] * uz + P[PXZ] * uy + P[PYZ] * ux);
}
template <class T>
MCG_HD T fneqrT(int i, const T P[6], T ux, T uy, T uz) {
    T a2 = T(0), a3 = T(0), t[7];
    for (int k = 0; k < 6; ++k) a2 += ((k == PXY || k == PYZ || k == PXZ) ? T(2) : T(1)) * h2T<T>(k, i) * P[k];
    fneqr_products<T>(P, ux, uy, uz, t);
    for (int k = 0; k < 7; ++k) a3 += h3T<T>(k, i) * t[k];
    return wT<T>(i) * (T(4.5) * a2 + T(13.5) * a3);
}
template <class T>
MCG_HD void fneqr_vjp(int i, const T P[6], T ux, T uy, T uz, T lam, T l_P[6], T l_u[3]) {
    const T w = wT<T>(i);
    for (int k = 0; k < 6; ++k)
        l_P[k] += lam * w * T(4.5) * ((k == PXY || k == PYZ || k == PXZ) ? T(2) : T(1)) * h2T<T>(k, i);
    T l_t[7];
    for (int k = 0; k < 7; ++k) l_t[k] = lam * w * T(13.5) * h3T<T>(k, i);
    t_vjp<T>(P, ux, uy, uz, l_t, l_P, l_u);
}

template <class T>
struct AdjExt {
    bool iface = false;  // per-node omega from tau(psi)
    T taur = T(0), taub = T(0);
    T grav = T(0);       // body force F_y = -grav rr
};

template <class T>
struct AdjNode {
    T rr, rb, rho, aR, aB, ux, uy, uz, usq;
    T P[6];
    T t[7];
    T tr;  // -(Pxx + Pyy + Pzz)/3
    T Fx, Fy, Fz, absF, hA, sB, s2, rc;
    T q2, q3, beta;
    T omega, sf, Fb, gy, g9, g3, e0;  // relaxation, body force, e0 with the force's u-only term
    AdjExt<T> ext;
    bool bulk, cut;
};

// Forward set-up of node y from its moments, its colour gradient and the
// parameters (A, beta, omega).
template <class T>
MCG_HD AdjNode<T> adj_setup(const T m[11], T Fx, T Fy, T Fz, T A, T beta, T omega, const AdjExt<T> &ext = AdjExt<T>{}) {
    AdjNode<T> n;
    n.ext = ext;
    n.rr = m[0];
    n.rb = m[1];
    n.rho = n.rr + n.rb;
    const T inv = T(1) / n.rho;
    n.aR = n.rr * inv;
    n.aB = n.rb * inv;
    n.ux = m[2];
    n.uy = m[3];
    n.uz = m[4];
    n.usq = n.ux * n.ux + n.uy * n.uy + n.uz * n.uz;
    for (int k = 0; k < 6; ++k) n.P[k] = m[5 + k];
    n.tr = -(n.P[PXX] + n.P[PYY] + n.P[PZZ]) / T(3);
    n.t[TXXY] = n.P[PXX] * n.uy + T(2) * n.P[PXY] * n.ux;
    n.t[TXXZ] = n.P[PXX] * n.uz + T(2) * n.P[PXZ] * n.ux;
    n.t[TXYY] = n.P[PYY] * n.ux + T(2) * n.P[PXY] * n.uy;
    n.t[TXZZ] = n.P[PZZ] * n.ux + T(2) * n.P[PXZ] * n.uz;
    n.t[TYYZ] = n.P[PYY] * n.uz + T(2) * n.P[PYZ] * n.uy;
    n.t[TYZZ] = n.P[PZZ] * n.uy + T(2) * n.P[PYZ] * n.uz;
    n.t[TXYZ] = T(2) * (n.P[PXY] * n.uz + n.P[PXZ] * n.uy + n.P[PYZ] * n.ux);
    if (ext.iface) {
        const T p = (n.rr - n.rb) / (n.rr + n.rb);
        omega = T(1) / (((T(1) + p) / T(2)) * ext.taur + ((T(1) - p) / T(2)) * ext.taub);
    }
    n.omega = omega;
    const T oms = T(1) - omega;
    n.q2 = T(4.5) * oms;
    n.q3 = T(13.5) * oms;
    n.beta = beta;
    n.sf = T(0.5) + T(0.5) * oms;
    n.Fb = -ext.grav * n.rr;
    n.gy = n.sf * n.Fb;
    n.g9 = T(9) * n.gy;
    n.g3 = T(1.5) * n.Fb;
    n.e0 = n.rho * (T(1) - T(1.5) * n.usq) - T(3) * n.gy * n.uy;
    n.bulk = T(4) * n.rr * n.rb * inv * inv <= T(1.0e-4);
    n.Fx = n.Fy = n.Fz = n.absF = n.hA = n.sB = n.s2 = n.rc = T(0);
    n.cut = true;
    if (!n.bulk) {
        n.Fx = Fx;
        n.Fy = Fy;
        n.Fz = Fz;
        n.hA = A / T(2);
        n.absF = sqrt(Fx * Fx + Fy * Fy + Fz * Fz);
        n.sB = n.hA * n.absF;
        n.cut = n.absF <= T(1.0e-4);
        if (!n.cut) {
            n.s2 = n.hA / n.absF;
            n.rc = beta * n.rr * n.rb / (n.rho * n.absF);
        }
    }
    return n;
}

template <class T>
MCG_HD T adj_a2(const AdjNode<T> &n, int i) {
    T a = n.tr;
    if (kCX[i] != 0) a += n.P[PXX];
    if (kCY[i] != 0) a += n.P[PYY];
    if (kCZ[i] != 0) a += n.P[PZZ];
    a += T(2 * kCX[i] * kCY[i]) * n.P[PXY];
    a += T(2 * kCY[i] * kCZ[i]) * n.P[PYZ];
    a += T(2 * kCX[i] * kCZ[i]) * n.P[PXZ];
    return a;
}

template <class T>
MCG_HD T adj_a3(const AdjNode<T> &n, int i) {
    T a = T(0);
    for (int k = 0; k < 7; ++k) a += h3T<T>(k, i) * n.t[k];
    return a;
}

// Forward post-collision populations of the node (used by the CPU adjoint's
// forward pass).
template <class T>
MCG_HD void adj_collide(const AdjNode<T> &n, T *out_r, T *out_b) {
    const T e0 = n.e0;
    {
        const T g0 = wT<T>(0) * (e0 + n.q2 * n.tr + T(1.25) * n.sB);
        out_r[0] = n.aR * g0;
        out_b[0] = n.aB * g0;
    }
    for (int k = 0; k < 13; ++k) {
        const int i = 2 * k + 1;
        const T cu = T(kCX[i]) * n.ux + T(kCY[i]) * n.uy + T(kCZ[i]) * n.uz;
        const T d = T(kCX[i]) * n.Fx + T(kCY[i]) * n.Fy + T(kCZ[i]) * n.Fz;
        const T E = e0 + T(4.5) * n.rho * cu * cu + n.q2 * adj_a2(n, i) + n.s2 * d * d - n.sB + T(kCY[i]) * n.g9 * cu;
        const T O = n.rho * cu * (T(3) - T(4.5) * n.usq) + T(4.5) * n.rho * cu * cu * cu + n.q3 * adj_a3(n, i) +
                    T(kCY[i]) * n.g3;
        const T D = n.rc * invcT<T>(i) * d;
        const T wi = wT<T>(i);
        const T gp = wi * (E + O), gm = wi * (E - O);
        out_r[i] = n.aR * gp + wi * D;
        out_b[i] = n.aB * gp - wi * D;
        out_r[i + 1] = n.aR * gm - wi * D;
        out_b[i + 1] = n.aB * gm + wi * D;
    }
}

// Adjoint of one node: lam_m[11] = d<mu, f*>/dm, lam_F[3] = d<mu, f*>/dF,
// lam_A and lam_beta the parameter sensitivities, for the adjoint mu_r, mu_b
// of the 54 post-collision populations.
template <class T>
struct AdjOut {
    T m[11];
    T F[3];
    T A, beta, grav;
};

// mu(i, r, b) yields the adjoint of the two colours of direction i; it is
// called once per direction, in pair order, so that callers can load it on
// demand.
template <class T, class MuFn>
MCG_HD AdjOut<T> adj_vjp_fn(const AdjNode<T> &n, MuFn &&mu) {
    T l_rho = 0, l_aR = 0, l_aB = 0, l_usq = 0, l_sB = 0, l_s2 = 0, l_rc = 0, l_tr = 0;
    T l_u[3] = {0, 0, 0}, l_P[6] = {0, 0, 0, 0, 0, 0}, l_t[7] = {0, 0, 0, 0, 0, 0, 0}, l_F[3] = {0, 0, 0};
    T l_e0 = 0, l_q2 = 0, l_q3 = 0, l_g9 = 0, l_g3 = 0;  // into the extension terms
    const T e0 = n.e0;
    // rest direction
    {
        const T g0 = wT<T>(0) * (e0 + n.q2 * n.tr + T(1.25) * n.sB);
        T mr0, mb0;
        mu(0, mr0, mb0);
        const T l_g0 = n.aR * mr0 + n.aB * mb0;
        l_aR += g0 * mr0;
        l_aB += g0 * mb0;
        const T l_E0 = wT<T>(0) * l_g0;
        l_e0 += l_E0;
        l_tr += n.q2 * l_E0;
        l_q2 += n.tr * l_E0;
        l_sB += T(1.25) * l_E0;
    }
    for (int k = 0; k < 13; ++k) {
        const int i = 2 * k + 1, j = i + 1;
        const T cx = T(kCX[i]), cy = T(kCY[i]), cz = T(kCZ[i]);
        const T cu = cx * n.ux + cy * n.uy + cz * n.uz;
        const T d = cx * n.Fx + cy * n.Fy + cz * n.Fz;
        const T a2 = adj_a2(n, i), a3 = adj_a3(n, i);
        const T E = e0 + T(4.5) * n.rho * cu * cu + n.q2 * a2 + n.s2 * d * d - n.sB + cy * n.g9 * cu;
        const T O = n.rho * cu * (T(3) - T(4.5) * n.usq) + T(4.5) * n.rho * cu * cu * cu + n.q3 * a3 + cy * n.g3;
        const T wi = wT<T>(i);
        const T gp = wi * (E + O), gm = wi * (E - O);
        T mri, mbi, mrj, mbj;
        mu(i, mri, mbi);
        mu(j, mrj, mbj);
        const T l_gp = n.aR * mri + n.aB * mbi;
        const T l_gm = n.aR * mrj + n.aB * mbj;
        l_aR += gp * mri + gm * mrj;
        l_aB += gp * mbi + gm * mbj;
        const T l_D = wi * (mri - mbi - mrj + mbj);
        const T l_E = wi * (l_gp + l_gm);
        const T l_O = wi * (l_gp - l_gm);
        // E
        l_e0 += l_E;
        l_rho += T(4.5) * cu * cu * l_E;
        T l_cu = T(9) * n.rho * cu * l_E + cy * n.g9 * l_E;
        l_g9 += cy * cu * l_E;
        const T l_a2 = n.q2 * l_E;
        l_q2 += a2 * l_E;
        l_s2 += d * d * l_E;
        T l_d = T(2) * n.s2 * d * l_E;
        l_sB += -l_E;
        // O
        l_rho += (cu * (T(3) - T(4.5) * n.usq) + T(4.5) * cu * cu * cu) * l_O;
        l_cu += n.rho * (T(3) - T(4.5) * n.usq + T(13.5) * cu * cu) * l_O;
        l_usq += -T(4.5) * n.rho * cu * l_O;
        const T l_a3 = n.q3 * l_O;
        l_q3 += a3 * l_O;
        l_g3 += cy * l_O;
        // D
        l_rc += invcT<T>(i) * d * l_D;
        l_d += n.rc * invcT<T>(i) * l_D;
        // c.u, c.F
        l_u[0] += cx * l_cu;
        l_u[1] += cy * l_cu;
        l_u[2] += cz * l_cu;
        l_F[0] += cx * l_d;
        l_F[1] += cy * l_d;
        l_F[2] += cz * l_d;
        // a2 = tr + sum_{c_a != 0} P_aa + 2 c_a c_b P_ab
        l_tr += l_a2;
        if (kCX[i] != 0) l_P[PXX] += l_a2;
        if (kCY[i] != 0) l_P[PYY] += l_a2;
        if (kCZ[i] != 0) l_P[PZZ] += l_a2;
        l_P[PXY] += T(2 * kCX[i] * kCY[i]) * l_a2;
        l_P[PYZ] += T(2 * kCY[i] * kCZ[i]) * l_a2;
        l_P[PXZ] += T(2 * kCX[i] * kCZ[i]) * l_a2;
        for (int q = 0; q < 7; ++q) l_t[q] += h3T<T>(q, i) * l_a3;
    }
    AdjOut<T> o;
    o.A = o.beta = o.grav = T(0);
    T l_rr = 0, l_rb = 0;
    // e0 = rho (1 - 1.5 usq) - 3 gy u_y
    l_rho += (T(1) - T(1.5) * n.usq) * l_e0;
    l_usq += -T(1.5) * n.rho * l_e0;
    T l_gy = -T(3) * n.uy * l_e0 + T(9) * l_g9;
    l_u[1] += -T(3) * n.gy * l_e0;
    // gy = sf Fb, g3 = 1.5 Fb, Fb = -grav rr, sf = 1 - omega/2, q2 = 4.5 (1 - omega), q3 = 13.5 (1 - omega)
    const T l_Fb = n.sf * l_gy + T(1.5) * l_g3;
    const T l_sf = n.Fb * l_gy;
    l_rr += -n.ext.grav * l_Fb;
    o.grav += -n.rr * l_Fb;
    if (n.ext.iface) {
        // omega = 1/tau(psi): d omega/d psi = -omega^2 (taur - taub)/2, psi = (rr - rb)/rho
        const T l_om = -T(0.5) * l_sf - T(4.5) * l_q2 - T(13.5) * l_q3;
        const T dpsi = -n.omega * n.omega * (n.ext.taur - n.ext.taub) / T(2) * l_om;
        l_rr += dpsi * T(2) * n.rb / (n.rho * n.rho);
        l_rb += -dpsi * T(2) * n.rr / (n.rho * n.rho);
    }
    // interface coefficients: sB = hA |F|, s2 = hA/|F|, rc = beta rr rb/(rho |F|)
    if (!n.bulk) {
        T l_absF = n.hA * l_sB;
        o.A += n.absF * l_sB / T(2);
        if (!n.cut) {
            o.A += l_s2 / (T(2) * n.absF);
            l_absF += -n.s2 / n.absF * l_s2;
            l_absF += -n.rc / n.absF * l_rc;
            o.beta += n.rr * n.rb / (n.rho * n.absF) * l_rc;
            // rc = beta rr rb / (rho |F|): d/drr = beta rb/(rho |F|) - rc/rho
            l_rr += (n.beta * n.rb / (n.rho * n.absF) - n.rc / n.rho) * l_rc;
            l_rb += (n.beta * n.rr / (n.rho * n.absF) - n.rc / n.rho) * l_rc;
        }
        if (n.absF > T(0)) {
            l_F[0] += n.Fx / n.absF * l_absF;
            l_F[1] += n.Fy / n.absF * l_absF;
            l_F[2] += n.Fz / n.absF * l_absF;
        }
    }
    // tr = -(Pxx + Pyy + Pzz)/3
    l_P[PXX] += -l_tr / T(3);
    l_P[PYY] += -l_tr / T(3);
    l_P[PZZ] += -l_tr / T(3);
    // third-order products
    t_vjp<T>(n.P, n.ux, n.uy, n.uz, l_t, l_P, l_u);
    // (the products' adjoint is added into l_P and l_u)
    const T *P = n.P;
    (void)P;
    // usq
    l_u[0] += T(2) * n.ux * l_usq;
    l_u[1] += T(2) * n.uy * l_usq;
    l_u[2] += T(2) * n.uz * l_usq;
    // aR = rr/rho, aB = rb/rho, rho = rr + rb
    const T inv = T(1) / n.rho;
    l_rr += l_aR * inv;
    l_rb += l_aB * inv;
    l_rho += -(n.rr * l_aR + n.rb * l_aB) * inv * inv;
    l_rr += l_rho;
    l_rb += l_rho;
    o.m[0] = l_rr;
    o.m[1] = l_rb;
    o.m[2] = l_u[0];
    o.m[3] = l_u[1];
    o.m[4] = l_u[2];
    for (int k = 0; k < 6; ++k) o.m[5 + k] = l_P[k];
    o.F[0] = l_F[0];
    o.F[1] = l_F[1];
    o.F[2] = l_F[2];
    return o;
}

template <class T>
MCG_HD AdjOut<T> adj_vjp(const AdjNode<T> &n, const T *mu_r, const T *mu_b) {
    return adj_vjp_fn<T>(n, [&](int i, T &r, T &b) {
        r = mu_r[i];
        b = mu_b[i];
    });
}

// Adjoint of the moments pass (lbm.cuh:82-165) at one node: from the moments
// m and their adjoint lam (11) to the adjoint of the 54 gathered populations.
//   rr = sum f_r, rb = sum f_b, u = sum g c / rho, Pi_ab = sum g H_ab - rho u_a u_b
template <class T>
MCG_HD void adj_moments(const T m[11], const T lam[11], T *l_fr, T *l_fb, T ghalf = T(0), T *dgrav = nullptr) {
    const T rho = m[0] + m[1];
    const T u[3] = {m[2], m[3], m[4]};
    // Pi_ab = S_ab - rho u_a u_b  ->  l_S = l_Pi, l_u, l_rho
    const T *lP = lam + 5;
    T l_u[3] = {lam[2], lam[3], lam[4]};
    T l_rho = 0;
    // diagonal
    l_u[0] += -T(2) * rho * u[0] * lP[PXX];
    l_u[1] += -T(2) * rho * u[1] * lP[PYY];
    l_u[2] += -T(2) * rho * u[2] * lP[PZZ];
    l_rho += -(u[0] * u[0] * lP[PXX] + u[1] * u[1] * lP[PYY] + u[2] * u[2] * lP[PZZ]);
    // off-diagonal
    l_u[0] += -rho * (u[1] * lP[PXY] + u[2] * lP[PXZ]);
    l_u[1] += -rho * (u[0] * lP[PXY] + u[2] * lP[PYZ]);
    l_u[2] += -rho * (u[1] * lP[PYZ] + u[0] * lP[PXZ]);
    l_rho += -(u[0] * u[1] * lP[PXY] + u[1] * u[2] * lP[PYZ] + u[0] * u[2] * lP[PXZ]);
    // u = j / rho
    const T inv = T(1) / rho;
    const T l_j[3] = {l_u[0] * inv, l_u[1] * inv, l_u[2] * inv};
    l_rho += -(l_u[0] * u[0] + l_u[1] * u[1] + l_u[2] * u[2]) * inv;
    // u_y = (j_y - ghalf rr)/rho, ghalf = grav/2: the explicit rr and grav dependence of the force shift
    const T l_rr = lam[0] + l_rho - ghalf * l_u[1] * inv, l_rb = lam[1] + l_rho;
    if (dgrav) *dgrav += -T(0.5) * m[0] * inv * l_u[1];
    for (int i = 0; i < Q; ++i) {
        T lg = T(kCX[i]) * l_j[0] + T(kCY[i]) * l_j[1] + T(kCZ[i]) * l_j[2];
        for (int k = 0; k < 6; ++k) lg += h2T<T>(k, i) * lP[k];
        l_fr[i] = l_rr + lg;
        l_fb[i] = l_rb + lg;
    }
}

}  // namespace mcg
