// mcg_oracle.cpp -- CPU restatement of the reference's D3Q27 colour-gradient
// lattice Boltzmann step.  TEST INFRASTRUCTURE ONLY: imported by tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker.
// Nothing in the product links it.
//
// It follows the reference's own data flow, not the GPU design: full-grid
// arrays with ghost layers, a moments pass (lbm.cuh:82-165) that writes the
// 11 moment arrays, an in-place push collide-stream pass (lbm.cuh:184-293),
// then the inlet (jet/boundary_conditions.cuh:17-85) and the convective outlet
// (jet/boundary_conditions.cuh:163-233), in the order of main.cu:54-66.
// Arithmetic is templated: T=double is the FP64 oracle, T=float is available
// for diagnostics.  OpenMP parallelises every sweep; the in-place push is
// race free because each population slot has at most one writer per step
// (checked by the write-set test).
//
// Parity status per function (pins live in tests/test_oracle_*.py):
//   lattice tables, feq, fneqr, Omega2, recolour, gradient: closed-form
//   invariants; moments pass: hand-set populations; full step: rest-state
//   fixed point, per-colour mass conservation (periodic), shear-wave decay
//   rate, Laplace law, write-set (quirks Q1/Q2), hand-computed inlet/outlet
//   node; radial statistics and TKE: brute-force recomputation.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/mcg.h"

namespace {

constexpr int Q = 27;

// D3Q27 velocities in the reference's order: 0 rest, 1-6 faces, 7-18 edges,
// 19-26 corners, opposite directions adjacent (stencil_ct.cuh:9-16).
// opposite direction (the directions 2k-1 and 2k are opposite)
constexpr int opposite(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

template <class T>
struct Lattice {
    T w[Q], B[Q], invc[Q];
    T Hxx[Q], Hxy[Q], Hyy[Q], Hyz[Q], Hzz[Q], Hxz[Q];
    T Hxxy[Q], Hxxz[Q], Hxyy[Q], Hxzz[Q], Hyyz[Q], Hyzz[Q], Hxyz[Q];
    T cs2, inv_cs2, inv_2cs2, inv_2cs4, inv_6cs6, inv_2cs6, cs4;
    Lattice() {
        // constants.cuh:45-53
        cs2 = T(1) / T(3);
        cs4 = cs2 * cs2;
        const T cs6 = cs4 * cs2;
        inv_cs2 = T(1) / cs2;
        inv_2cs2 = T(1) / (T(2) * cs2);
        inv_2cs4 = T(1) / (T(2) * cs4);
        inv_6cs6 = T(1) / (T(6) * cs6);
        inv_2cs6 = T(1) / (T(2) * cs6);
        for (int i = 0; i < Q; ++i) {
            const int s = CX[i] * CX[i] + CY[i] * CY[i] + CZ[i] * CZ[i];
            // weights stencil_ct.cuh:42-65, perturbation coefficients :67-78
            w[i] = s == 0 ? T(8) / T(27) : s == 1 ? T(2) / T(27) : s == 2 ? T(1) / T(54) : T(1) / T(216);
            B[i] = i == 0 ? T(-10) / T(27) : w[i];
            // 1/|c_i| stencil_ct.cuh:168-190
            invc[i] = s == 0 ? T(0) : T(1) / std::sqrt(T(s));
            const T cx = T(CX[i]), cy = T(CY[i]), cz = T(CZ[i]);
            // Hermite polynomials stencil_ct.cuh:85-164
            Hxx[i] = cx * cx - cs2;
            Hxy[i] = cx * cy;
            Hyy[i] = cy * cy - cs2;
            Hyz[i] = cy * cz;
            Hzz[i] = cz * cz - cs2;
            Hxz[i] = cx * cz;
            Hxxy[i] = cx * cx * cy - cs2 * cy;
            Hxxz[i] = cx * cx * cz - cs2 * cz;
            Hxyy[i] = cx * cy * cy - cs2 * cx;
            Hxzz[i] = cx * cz * cz - cs2 * cx;
            Hyyz[i] = cy * cy * cz - cs2 * cz;
            Hyzz[i] = cy * cz * cz - cs2 * cy;
            Hxyz[i] = cx * cy * cz;
        }
    }
};

// Third-order Hermite equilibrium, lbm.cuh:15-35.
template <class T>
inline T feq(const Lattice<T> &L, int i, T rho, T ux, T uy, T uz) {
    const T cu = ux * T(CX[i]) + uy * T(CY[i]) + uz * T(CZ[i]);
    const T usq = ux * ux + uy * uy + uz * uz;
    const T A2 = cu * cu * L.inv_2cs4 - usq * L.inv_2cs2 + cu * L.inv_cs2;
    const T A3 = cu * (cu * cu * L.inv_6cs6 - usq * L.inv_2cs4);
    return L.w[i] * rho * (T(1) + A2 + A3);
}

// Regularised non-equilibrium part rebuilt from Pi^neq and u, lbm.cuh:37-78.
template <class T>
inline T fneqr(const Lattice<T> &L, int i, const T P[6], T ux, T uy, T uz) {
    // P = {xx, xy, yy, yz, zz, xz}
    const T A2 = (P[0] * L.Hxx[i] + T(2) * P[1] * L.Hxy[i] + P[2] * L.Hyy[i] + T(2) * P[3] * L.Hyz[i] +
                  P[4] * L.Hzz[i] + T(2) * P[5] * L.Hxz[i]) * L.inv_2cs4;
    const T a3xxy = P[0] * uy + T(2) * P[1] * ux;
    const T a3xxz = P[0] * uz + T(2) * P[5] * ux;
    const T a3xyy = P[2] * ux + T(2) * P[1] * uy;
    const T a3xzz = P[4] * ux + T(2) * P[5] * uz;
    const T a3yyz = P[2] * uz + T(2) * P[3] * uy;
    const T a3yzz = P[4] * uy + T(2) * P[3] * uz;
    const T a3xyz = P[1] * uz + P[5] * uy + P[3] * ux;
    const T A3 = (a3xxy * L.Hxxy[i] + a3xxz * L.Hxxz[i] + a3xyy * L.Hxyy[i] + a3xzz * L.Hxzz[i] +
                  a3yyz * L.Hyyz[i] + a3yzz * L.Hyzz[i] + T(2) * a3xyz * L.Hxyz[i]) * L.inv_2cs6;
    return L.w[i] * (A2 + A3);
}

// (F.c_i)^2/|F|^2 with the |F|<=1e-4 cut, forces.cuh:50-75.
template <class T>
inline T cos2rule(int i, T Fx, T Fy, T Fz, T absF) {
    if (i == 0) return T(0);
    if (absF <= T(1.0e-4)) return T(0);
    const T d = Fx * T(CX[i]) + Fy * T(CY[i]) + Fz * T(CZ[i]);
    return (d * d) / (absF * absF);
}

// Perturbation operator, collisionOperators.cuh:15-24.
template <class T>
inline T omega2(const Lattice<T> &L, int i, T Fx, T Fy, T Fz, T absF, T A) {
    return (A / T(2)) * absF * (L.w[i] * cos2rule<T>(i, Fx, Fy, Fz, absF) - L.B[i]);
}

// cos of the angle between F and c_i, collisionOperators.cuh:28-49.
template <class T>
inline T cosphi(const Lattice<T> &L, int i, T Fx, T Fy, T Fz, T absF) {
    if (i == 0) return T(0);
    if (absF <= T(1.0e-4)) return T(0);
    const T d = Fx * T(CX[i]) + Fy * T(CY[i]) + Fz * T(CZ[i]);
    return d * (T(1) / absF) * L.invc[i];
}

// Latva-Kokko recolouring increment, collisionOperators.cuh:51-69.
template <class T>
inline T recolor(const Lattice<T> &L, int i, T rr, T rb, T Fx, T Fy, T Fz, T absF, T beta) {
    if (i == 0) return T(0);
    const T rho = rr + rb;
    const T fe = L.w[i] * rho;
    const T c = cosphi<T>(L, i, Fx, Fy, Fz, absF);
    const T coeff = beta * ((rr * rb) / (rho * rho));
    return coeff * c * fe;
}

// phase field, utilities/mathUtilities.cuh:24-28
template <class T>
inline T psi(T self, T other) { return (self - other) / (self + other); }

template <class T>
inline T clamp01(T x) { return std::fmin(std::fmax(x, T(0)), T(1)); }

struct OracleBase {
    virtual ~OracleBase() {}
    virtual int init(int kind) = 0;
    virtual int step(int n) = 0;
    virtual void get_fields(double *rr, double *rb, double *ux, double *uy, double *uz) = 0;
    virtual void get_moments(double *out) = 0;
    virtual void get_populations(double *fr, double *fb) = 0;
    virtual void set_populations(const double *fr, const double *fb) = 0;
    virtual void set_densities(const double *rr, const double *rb) = 0;
    virtual void get_radial(double *, double *, double *, double *, double *, uint64_t *, double *, double *) = 0;
    virtual void stats_info(int32_t *, int64_t *, int64_t *) = 0;
    virtual double tke() = 0;
    virtual void global_sums(double *) = 0;
    virtual int writeset(uint8_t *cr, uint8_t *cb) = 0;
    virtual void gradient(int x, int y, int z, double F[3]) = 0;
    virtual int64_t get_step() = 0;
    virtual void set_inlet_offset(const double *du) = 0;
    virtual double uy_dot(const double *w) = 0;
    virtual void branch_masks(uint8_t *out) = 0;
};

template <class T>
struct Oracle final : OracleBase {
    mcg_case c;
    Lattice<T> L;
    int NX, NY, NZ;
    size_t N;
    int NR;
    std::vector<T> fr, fb;
    std::vector<T> rhor, rhob, ux, uy, uz, Pxx, Pxy, Pyy, Pyz, Pzz, Pxz;
    std::vector<double> s_uy, s_uy2, s_ur, s_ur2, s_uruy, s_phi, s_phi2;
    std::vector<uint64_t> s_cnt;
    int64_t step_ = 0;
    std::vector<double> inlet_du;  // optional per-column inlet velocity offset, 3 x NX*NZ
    bool started_ = false;
    int64_t start_step_ = 0, samples_ = 0;
    bool ready_ = false;
    // write-set counting mode
    uint8_t *cnt_r = nullptr, *cnt_b = nullptr;
    T nu, taub_, taur_, beta_, sig4;

    explicit Oracle(const mcg_case &cc) : c(cc) {
        NX = c.NX; NY = c.NY; NZ = c.NZ;
        N = size_t(NX) * NY * NZ;
        NR = NX / 2;  // constants.cuh:20
        fr.assign(size_t(Q) * N, T(0));
        fb.assign(size_t(Q) * N, T(0));
        for (auto *v : {&rhor, &rhob, &ux, &uy, &uz, &Pxx, &Pxy, &Pyy, &Pyz, &Pzz, &Pxz}) v->assign(N, T(0));
        const size_t nb = size_t(NY) * NR;
        for (auto *v : {&s_uy, &s_uy2, &s_ur, &s_ur2, &s_uruy, &s_phi, &s_phi2}) v->assign(nb, 0.0);
        s_cnt.assign(nb, 0);
        taub_ = T(c.taub);
        taur_ = T(c.taur);
        nu = (taub_ - T(0.5)) * L.cs2;  // inverse of constants.cuh:64
        beta_ = T(c.beta_recolor);
        sig4 = T(c.sigma) / (T(4) * L.cs4);  // constants.cuh:76
    }

    inline size_t idx(int x, int y, int z) const { return size_t(x) + size_t(NX) * (size_t(y) + size_t(NY) * size_t(z)); }
    inline size_t fidx(size_t id, int i) const { return size_t(i) * N + id; }
    inline bool interior(int x, int y, int z) const {
        return x >= 1 && x <= NX - 2 && y >= 1 && y <= NY - 2 && z >= 1 && z <= NZ - 2;
    }
    // periodic wrap with ghost layers, utilities/bounds.cuh:23-60
    inline int wrapx(int x) const { return x == 0 ? NX - 2 : (x == NX - 1 ? 1 : x); }
    inline int wrapy(int y) const { return y == 0 ? NY - 2 : (y == NY - 1 ? 1 : y); }
    inline int wrapz(int z) const { return z == 0 ? NZ - 2 : (z == NZ - 1 ? 1 : z); }

    // sponge-layer relaxation rate, utilities/mathUtilities.cuh:49-90
    T omega_sponge(int y) const {
        const int yend = NY - 2;
        const int ystart = yend - c.sponge_cells + 1;
        const T inv_width = T(1) / T(yend - ystart);
        const T s = clamp01<T>((T(y) - T(ystart)) * inv_width);
        const T nus = nu * (T(1) + T(c.sponge_K) * s * s * s);
        const T tau = T(0.5) + nus * L.inv_cs2;
        return T(1) / tau;
    }
    // relaxation time used for A and, in MCG_VISC_INTERFACE mode, for omega (forces.cuh:77-105)
    T tau_node(T rr, T rb) const {
        if (c.viscosity_mode == MCG_VISC_INTERFACE) {
            const T p = psi<T>(rr, rb);
            return ((T(1) + p) / T(2)) * taur_ + ((T(1) - p) / T(2)) * taub_;
        }
        return taub_;
    }
    T omega_node(int y, T rr, T rb) const {
        if (c.viscosity_mode == MCG_VISC_INTERFACE) return T(1) / tau_node(rr, rb);
        return omega_sponge(y);
    }
    T A_node(T rr, T rb) const {
        if (c.A_override > 0) return T(c.A_override);
        return sig4 / tau_node(rr, rb);  // forces.cuh:87-90,102-103
    }
    bool is_jet(int x, int z) const {  // utilities/mathUtilities.cuh:14-21
        const T dx = T(x) - T(c.jet_x0), dz = T(z) - T(c.jet_z0);
        const T R = T(c.jet_radius);
        return dx * dx + dz * dz <= R * R;
    }

    inline void put(std::vector<T> &f, uint8_t *cnt, size_t slot, T v) {
        if (cnt) {
            cnt[slot]++;
        } else {
            f[slot] = v;
        }
    }

    void reset_stats() {
        for (auto *v : {&s_uy, &s_uy2, &s_ur, &s_ur2, &s_uruy, &s_phi, &s_phi2}) std::fill(v->begin(), v->end(), 0.0);
        std::fill(s_cnt.begin(), s_cnt.end(), 0);
        started_ = false;
        start_step_ = 0;
        samples_ = 0;
    }

    int init(int kind) override {
        for (auto *v : {&ux, &uy, &uz, &Pxx, &Pxy, &Pyy, &Pyz, &Pzz, &Pxz}) std::fill(v->begin(), v->end(), T(0));
        // all-node initialisation, jet/jetIn.cuh:13-35 / bubble/bubbleIn.cuh:10-39
#pragma omp parallel for schedule(static)
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y)
                for (int x = 0; x < NX; ++x) {
                    const size_t id = idx(x, y, z);
                    T rr, rb;
                    if (kind == 0) {
                        rr = T(c.rhor0);
                        rb = T(0);
                    } else if (kind == 1) {
                        const T dx = T(x) - T(c.bubble_x0), dy = T(y) - T(c.bubble_y0), dz = T(z) - T(c.bubble_z0);
                        const T R = T(c.bubble_radius);
                        const T in = (dx * dx + dy * dy + dz * dz) <= R * R ? T(1) : T(0);
                        rr = (T(1) - in) * T(c.rhor0);
                        rb = in * T(c.rhob0);
                    } else {
                        const T dx = T(x) - T(c.bubble_x0), dy = T(y) - T(c.bubble_y0), dz = T(z) - T(c.bubble_z0);
                        const T r = std::sqrt(dx * dx + dy * dy + dz * dz);
                        const T frac = (T(1) + std::tanh((T(c.bubble_radius) - r) / T(c.interface_width))) / T(2);
                        rr = (T(1) - frac) * T(c.rhor0);
                        rb = frac * T(c.rhob0);
                    }
                    rhor[id] = rr;
                    rhob[id] = rb;
                    for (int i = 0; i < Q; ++i) {
                        if (kind == 0) {
                            fr[fidx(id, i)] = feq<T>(L, i, rr, T(0), T(0), T(0));
                            fb[fidx(id, i)] = feq<T>(L, i, rb, T(0), T(0), T(0));
                        } else {
                            fr[fidx(id, i)] = L.w[i] * rr;
                            fb[fidx(id, i)] = L.w[i] * rb;
                        }
                    }
                }
        if (kind == 0 && c.y_mode == MCG_Y_JET) {
            // jet mask on the inlet ghost plane, jet/jetIn.cuh:37-62
            for (int z = 1; z <= NZ - 2; ++z)
                for (int x = 1; x <= NX - 2; ++x) {
                    const size_t id = idx(x, 0, z);
                    const T j = is_jet(x, z) ? T(1) : T(0);
                    rhor[id] = (T(1) - j) * T(c.rhor0);
                    rhob[id] = j * T(c.rhob0);
                    for (int i = 0; i < Q; ++i) {
                        fr[fidx(id, i)] = feq<T>(L, i, rhor[id], T(0), T(0), T(0));
                        fb[fidx(id, i)] = feq<T>(L, i, rhob[id], T(0), T(0), T(0));
                    }
                }
        }
        step_ = 0;
        reset_stats();
        ready_ = true;
        return MCG_OK;
    }

    // Moments pass, lbm.cuh:82-165 (interior nodes only, kernels.cu:60-76)
    void mfields() {
#pragma omp parallel for schedule(static)
        for (int z = 1; z <= NZ - 2; ++z)
            for (int y = 1; y <= NY - 2; ++y)
                for (int x = 1; x <= NX - 2; ++x) {
                    const size_t id = idx(x, y, z);
                    T sr = 0, sb = 0, jx = 0, jy = 0, jz = 0, axx = 0, axy = 0, ayy = 0, ayz = 0, azz = 0, axz = 0;
                    for (int i = 0; i < Q; ++i) {
                        const T a = fr[fidx(id, i)], b = fb[fidx(id, i)];
                        sr += a;
                        sb += b;
                        const T g = a + b;
                        jx += g * T(CX[i]);
                        jy += g * T(CY[i]);
                        jz += g * T(CZ[i]);
                        axx += g * L.Hxx[i];
                        axy += g * L.Hxy[i];
                        ayy += g * L.Hyy[i];
                        ayz += g * L.Hyz[i];
                        azz += g * L.Hzz[i];
                        axz += g * L.Hxz[i];
                    }
                    rhor[id] = sr;
                    rhob[id] = sb;
                    const T rho = sr + sb;
                    const T inv = T(1) / rho;
                    // Guo forcing (an extension): u = (j + F/2)/rho, F = (0, -gravity rho_r, 0)
                    const T vx = jx * inv, vy = (jy - T(0.5) * T(c.gravity) * sr) * inv, vz = jz * inv;
                    ux[id] = vx;
                    uy[id] = vy;
                    uz[id] = vz;
                    Pxx[id] = axx - rho * vx * vx;
                    Pxy[id] = axy - rho * vx * vy;
                    Pyy[id] = ayy - rho * vy * vy;
                    Pyz[id] = ayz - rho * vy * vz;
                    Pzz[id] = azz - rho * vz * vz;
                    Pxz[id] = axz - rho * vx * vz;
                }
    }

    // colour gradient, forces.cuh:13-43 (neighbours read at flat offsets;
    // wrap_gradient=1 applies the periodic wrap of the streaming step instead)
    void force(int x, int y, int z, T &Fx, T &Fy, T &Fz) const {
        T sx = 0, sy = 0, sz = 0;
        for (int i = 1; i < Q; ++i) {
            int xn = x + CX[i], yn = y + CY[i], zn = z + CZ[i];
            if (c.wrap_gradient) {  // along the periodic axes
                if (!c.wall_xz) {
                    xn = wrapx(xn);
                    zn = wrapz(zn);
                }
                if (c.y_mode == MCG_Y_PERIODIC) yn = wrapy(yn);
            }
            const size_t n = idx(xn, yn, zn);
            const T p = psi<T>(rhor[n], rhob[n]);
            sx += L.w[i] * p * T(CX[i]);
            sy += L.w[i] * p * T(CY[i]);
            sz += L.w[i] * p * T(CZ[i]);
        }
        Fx = L.inv_cs2 * sx;
        Fy = L.inv_cs2 * sy;
        Fz = L.inv_cs2 * sz;
    }

    // The branches the collision of the last step took, per interior node (x
    // fastest): bit 0 = interface (I > 1e-4, lbm.cuh:214-216), bit 1 = |F| >
    // 1e-4 at an interface node (the cut of forces.cuh:61-64 and
    // collisionOperators.cuh:36-39 not taken).  The same tests as collistream,
    // on the moments it used.
    void branch_masks(uint8_t *out) override {
        const size_t ix = size_t(NX - 2), iy = size_t(NY - 2);
#pragma omp parallel for schedule(static)
        for (int z = 1; z <= NZ - 2; ++z)
            for (int y = 1; y <= NY - 2; ++y)
                for (int x = 1; x <= NX - 2; ++x) {
                    const size_t id = idx(x, y, z);
                    const T rr = rhor[id], rb = rhob[id];
                    const T invrT = T(1) / (rr + rb);
                    const T I = T(4) * rr * rb * invrT * invrT;
                    uint8_t m = 0;
                    if (!(I <= T(1e-4))) {
                        T Fx, Fy, Fz;
                        force(x, y, z, Fx, Fy, Fz);
                        const T absF = std::sqrt(Fx * Fx + Fy * Fy + Fz * Fz);
                        m = absF <= T(1.0e-4) ? 1 : 3;
                    }
                    out[size_t(x - 1) + ix * (size_t(y - 1) + iy * size_t(z - 1))] = m;
                }
    }

    // Collision + perturbation + recolouring + push streaming, lbm.cuh:184-293
    void collistream() {
#pragma omp parallel for schedule(static)
        for (int z = 1; z <= NZ - 2; ++z)
            for (int y = 1; y <= NY - 2; ++y)
                for (int x = 1; x <= NX - 2; ++x) {
                    const size_t id = idx(x, y, z);
                    const T rr = rhor[id], rb = rhob[id];
                    const T rT = rr + rb;
                    const T invrT = T(1) / rT;
                    const T aR = rr * invrT, aB = rb * invrT;
                    const T vx = ux[id], vy = uy[id], vz = uz[id];
                    const T P[6] = {Pxx[id], Pxy[id], Pyy[id], Pyz[id], Pzz[id], Pxz[id]};
                    const T om = omega_node(y, rr, rb);
                    const T oms = T(1) - om;
                    const T I = T(4) * rr * rb * invrT * invrT;
                    const bool bulk = I <= T(1e-4);
                    T Fx = 0, Fy = 0, Fz = 0, absF = 0, A = 0;
                    if (!bulk) {
                        force(x, y, z, Fx, Fy, Fz);
                        A = A_node(rr, rb);
                        absF = std::sqrt(Fx * Fx + Fy * Fy + Fz * Fz);
                    }
                    // body force on red (an extension): Guo source (1 - omega/2) w_i [3 (c_i - u) +
                    // 9 (c_i.u) c_i].F plus the (1 - omega) share of the first-order non-equilibrium
                    // moment -F/2 that the regularisation drops: odd part 1.5 w_i c_i.F in total
                    const T Fb = -T(c.gravity) * rr, sf = T(1) - om / T(2);
                    for (int i = 0; i < Q; ++i) {
                        T g = feq<T>(L, i, rT, vx, vy, vz) + oms * fneqr<T>(L, i, P, vx, vy, vz);
                        if (c.gravity != 0.0) {
                            const T cu = T(CX[i]) * vx + T(CY[i]) * vy + T(CZ[i]) * vz;
                            g = g + L.w[i] * (sf * (T(9) * cu * T(CY[i]) - T(3) * vy) * Fb + T(1.5) * T(CY[i]) * Fb);
                        }
                        T d = 0;
                        if (!bulk) {
                            g = g + omega2<T>(L, i, Fx, Fy, Fz, absF, A);
                            d = recolor<T>(L, i, rr, rb, Fx, Fy, Fz, absF, beta_);
                        }
                        size_t s;
                        if (c.wall_xz && (x + CX[i] < 1 || x + CX[i] > NX - 2 || z + CZ[i] < 1 || z + CZ[i] > NZ - 2)) {
                            s = fidx(id, opposite(i));  // half-way bounce-back at an x/z wall (an extension)
                        } else {
                            const int xn = wrapx(x + CX[i]);
                            const int zn = wrapz(z + CZ[i]);
                            int yn = y + CY[i];
                            if (c.y_mode == MCG_Y_PERIODIC) yn = wrapy(yn);
                            s = fidx(idx(xn, yn, zn), i);
                        }
                        if (bulk) {
                            put(fr, cnt_r, s, aR * g);
                            put(fb, cnt_b, s, aB * g);
                        } else {
                            put(fr, cnt_r, s, aR * g + d);
                            put(fb, cnt_b, s, aB * g - d);
                        }
                    }
                }
    }

    // SplitMix64 (Steele, Lea and Flood 2014; Vigna's reference constants), the
    // counter-based generator of the MCG_INLET_TANH perturbation
    static uint64_t splitmix64(uint64_t x) {
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    }
    // uniform in [0,1) with 24 bits: number k of column (x,z) at step n
    double inlet_xi(int64_t n, int x, int z, int k) const {
        const uint64_t key = splitmix64(splitmix64(splitmix64(uint64_t(uint32_t(c.perturbation_seed))) ^ uint64_t(n)) ^
                                        (uint64_t(z) * uint64_t(NX) + uint64_t(x)));
        return double(splitmix64(key + uint64_t(k)) >> 40) / 16777216.0;
    }
    // Inlet velocity: (0, U, 0) for the reference's sharp profile
    // (jet/boundary_conditions.cuh:45-47); for MCG_INLET_TANH (an extension)
    // u_y = U (1 - tanh((r - R)/w))/2 and u_x, u_z = amp U (2 xi - 1)
    void inlet_velocity(int x, int z, T &ex, T &ey, T &ez) const {
        ex = T(0);
        ey = T(c.jet_velocity);
        ez = T(0);
        if (c.inlet_profile == MCG_INLET_TANH) {
            const T dx = T(x) - T(c.jet_x0), dz = T(z) - T(c.jet_z0);
            const T r = std::sqrt(dx * dx + dz * dz);
            ey = T(c.jet_velocity) * (T(0.5) * (T(1) - std::tanh((r - T(c.jet_radius)) / T(c.interface_width))));
            const T a = T(c.perturbation_amp) * T(c.jet_velocity);
            ex = a * (T(2) * T(inlet_xi(step_, x, z, 1)) - T(1));
            ez = a * (T(2) * T(inlet_xi(step_, x, z, 2)) - T(1));
        }
        if (!inlet_du.empty()) {
            const size_t nxz = size_t(NX) * NZ, o = size_t(z) * NX + x;
            ex += T(inlet_du[o]);
            ey += T(inlet_du[nxz + o]);
            ez += T(inlet_du[2 * nxz + o]);
        }
    }
    // Offset added to the inlet velocity of every jet column (the control of
    // the inverse jet problem; an extension).  Null clears it.
    void set_inlet_offset(const double *du) override {
        if (!du) {
            inlet_du.clear();
            return;
        }
        inlet_du.assign(du, du + 3 * size_t(NX) * NZ);
    }

    // Inlet, jet/boundary_conditions.cuh:17-85 (inner (x,z) only, kernels.cu:98-114)
    void inlet() {
        for (int z = 1; z <= NZ - 2; ++z)
            for (int x = 1; x <= NX - 2; ++x) {
                const size_t idB = idx(x, 0, z), idF = idx(x, 1, z);
                const bool j = is_jet(x, z);
                rhor[idB] = (T(1) - (j ? T(1) : T(0))) * T(c.rhor0);
                rhob[idB] = (j ? T(1) : T(0)) * T(c.rhob0);
                if (!j) continue;
                const T rT = rhor[idB] + rhob[idB];
                const T P[6] = {Pxx[idF], Pxy[idF], Pyy[idF], Pyz[idF], Pzz[idF], Pxz[idF]};
                const T uxf = ux[idF], uyf = uy[idF], uzf = uz[idF];
                const T aR = rhor[idB] * (T(1) / rT), aB = rhob[idB] * (T(1) / rT);
                T om = T(1) / taub_;  // omegab, constants.cuh:67
                if (c.viscosity_mode == MCG_VISC_INTERFACE) om = T(1) / tau_node(rhor[idB], rhob[idB]);
                const T oms = T(1) - om;
                T ex, ey, ez;
                inlet_velocity(x, z, ex, ey, ez);
                for (int i = 0; i < Q; ++i) {
                    if (CY[i] != 1) continue;
                    const size_t t = idx(x + CX[i], 1, z + CZ[i]);  // offset_xz, no wrap
                    const T g = feq<T>(L, i, rT, ex, ey, ez) + oms * fneqr<T>(L, i, P, uxf, uyf, uzf);
                    put(fr, cnt_r, fidx(t, i), aR * g);
                    put(fb, cnt_b, fidx(t, i), aB * g);
                }
            }
    }

    // Convective outlet, jet/boundary_conditions.cuh:163-233
    void outlet() {
        const int yB = NY - 1, yF = NY - 2;
        for (int z = 1; z <= NZ - 2; ++z)
            for (int x = 1; x <= NX - 2; ++x) {
                const size_t idB = idx(x, yB, z), idF = idx(x, yF, z);
                const T rrF = rhor[idF], rbF = rhob[idF], uxF = ux[idF], uyF = uy[idF], uzF = uz[idF];
                const T uc = clamp01<T>(uyF);
                // phiB_old - uc (phiB_old - phiF), jet/boundary_conditions.cuh:87-93
                const T rrB = rhor[idB] - uc * (rhor[idB] - rrF);
                const T rbB = rhob[idB] - uc * (rhob[idB] - rbF);
                const T uxB = ux[idB] - uc * (ux[idB] - uxF);
                const T uyB = uy[idB] - uc * (uy[idB] - uyF);
                const T uzB = uz[idB] - uc * (uz[idB] - uzF);
                rhor[idB] = rrB;
                rhob[idB] = rbB;
                ux[idB] = uxB;
                uy[idB] = uyB;
                uz[idB] = uzB;
                const T rhoT = rrB + rbB;
                const T inv = T(1) / rhoT;
                const T P[6] = {Pxx[idF], Pxy[idF], Pyy[idF], Pyz[idF], Pzz[idF], Pxz[idF]};
                const T aR = rrB * inv, aB = rbB * inv;
                const T om = omega_node(yF, rrB, rbB);
                const T oms = T(1) - om;
                for (int i = 0; i < Q; ++i) {
                    if (CY[i] != -1) continue;
                    const size_t t = idx(x + CX[i], yF, z + CZ[i]);
                    const T g = feq<T>(L, i, rhoT, uxB, uyB, uzB) + oms * fneqr<T>(L, i, P, uxF, uyF, uzF);
                    put(fr, cnt_r, fidx(t, i), aR * g);
                    put(fb, cnt_b, fidx(t, i), aB * g);
                }
            }
    }

    // Radial moments, kernels.cu:188-233, plus psi moments (extension)
    void accumulate_radial() {
#pragma omp parallel for schedule(static)
        for (int y = 1; y <= NY - 2; ++y)
            for (int z = 1; z <= NZ - 2; ++z)
                for (int x = 1; x <= NX - 2; ++x) {
                    const double dx = double(x) - c.jet_x0, dz = double(z) - c.jet_z0;
                    const double r = std::sqrt(dx * dx + dz * dz);
                    const long rb = long(r);
                    if (rb >= NR) continue;
                    const size_t id = idx(x, y, z);
                    const size_t mid = size_t(y) * NR + size_t(rb);
                    const double vy = double(uy[id]);
                    double vr = 0;
                    if (r > 1.0e-12) vr = (double(ux[id]) * dx + double(uz[id]) * dz) / r;
                    const double p = double(psi<T>(rhor[id], rhob[id]));
                    s_uy[mid] += vy;
                    s_uy2[mid] += vy * vy;
                    s_ur[mid] += vr;
                    s_ur2[mid] += vr * vr;
                    s_uruy[mid] += vr * vy;
                    s_phi[mid] += p;
                    s_phi2[mid] += p * p;
                    s_cnt[mid] += 1;
                }
    }

    int one_step(bool count_only) {
        mfields();
        // update_radial_statistics_start, meanF/meanFieldsRT.cuh:173-189
        if (!started_ && step_ >= c.NSTATS_START_STEP) {
            started_ = true;
            start_step_ = step_;
            samples_ = 0;
        }
        if (started_ && !count_only) {
            const int64_t every = c.NSTATS_SAMPLE > 0 ? c.NSTATS_SAMPLE : 1;
            if ((step_ - start_step_) % every == 0) {
                accumulate_radial();
                ++samples_;
            }
        }
        collistream();
        if (c.y_mode == MCG_Y_JET) {
            inlet();
            outlet();
        }
        ++step_;
        return MCG_OK;
    }

    int step(int n) override {
        if (!ready_) return MCG_E_STATE;
        for (int k = 0; k < n; ++k) one_step(false);
        return MCG_OK;
    }

    int writeset(uint8_t *cr, uint8_t *cb) override {
        if (!ready_) return MCG_E_STATE;
        cnt_r = cr;
        cnt_b = cb;
        // counting mode: moments are still computed (the boundary passes need them)
        mfields();
        collistream();
        if (c.y_mode == MCG_Y_JET) {
            // inlet/outlet also update ghost moments; keep a copy so the state is unchanged
            std::vector<T> r0 = rhor, b0 = rhob, x0 = ux, y0 = uy, z0 = uz;
            inlet();
            outlet();
            rhor = r0; rhob = b0; ux = x0; uy = y0; uz = z0;
        }
        cnt_r = cnt_b = nullptr;
        return MCG_OK;
    }

    void get_fields(double *a, double *b, double *x, double *y, double *z) override {
        for (size_t i = 0; i < N; ++i) {
            a[i] = double(rhor[i]);
            b[i] = double(rhob[i]);
            x[i] = double(ux[i]);
            y[i] = double(uy[i]);
            z[i] = double(uz[i]);
        }
    }
    // sum_x w(x) u_y(x) over the grid, u_y as get_fields returns it (the
    // moments pass of the last step); nodes with w = 0 are skipped
    double uy_dot(const double *w) override {
        double acc = 0.0;
        for (size_t i = 0; i < N; ++i)
            if (w[i] != 0.0) acc += w[i] * double(uy[i]);
        return acc;
    }
    void get_moments(double *out) override {
        const std::vector<T> *arr[11] = {&rhor, &rhob, &ux, &uy, &uz, &Pxx, &Pxy, &Pyy, &Pyz, &Pzz, &Pxz};
        for (int k = 0; k < 11; ++k)
            for (size_t i = 0; i < N; ++i) out[size_t(k) * N + i] = double((*arr[k])[i]);
    }
    void get_populations(double *a, double *b) override {
        for (size_t i = 0; i < size_t(Q) * N; ++i) {
            a[i] = double(fr[i]);
            b[i] = double(fb[i]);
        }
    }
    void set_populations(const double *a, const double *b) override {
        for (size_t i = 0; i < size_t(Q) * N; ++i) {
            fr[i] = T(a[i]);
            fb[i] = T(b[i]);
        }
    }
    void set_densities(const double *a, const double *b) override {
        for (size_t i = 0; i < N; ++i) {
            rhor[i] = T(a[i]);
            rhob[i] = T(b[i]);
        }
    }
    void get_radial(double *a, double *b, double *cc, double *d, double *e, uint64_t *n, double *p, double *p2) override {
        const size_t nb = size_t(NY) * NR;
        if (a) std::memcpy(a, s_uy.data(), nb * 8);
        if (b) std::memcpy(b, s_uy2.data(), nb * 8);
        if (cc) std::memcpy(cc, s_ur.data(), nb * 8);
        if (d) std::memcpy(d, s_ur2.data(), nb * 8);
        if (e) std::memcpy(e, s_uruy.data(), nb * 8);
        if (n) std::memcpy(n, s_cnt.data(), nb * 8);
        if (p) std::memcpy(p, s_phi.data(), nb * 8);
        if (p2) std::memcpy(p2, s_phi2.data(), nb * 8);
    }
    void stats_info(int32_t *st, int64_t *ss, int64_t *ns) override {
        if (st) *st = started_ ? 1 : 0;
        if (ss) *ss = start_step_;
        if (ns) *ns = samples_;
    }
    // compute_total_tke, kernels.cu:136-152 and meanF/meanFields.cuh:10-19 (FP64 sum)
    // Sums are formed per z plane and then added in plane order, so the
    // result does not depend on the OpenMP thread count or timing.
    double tke() override {
        std::vector<double> part(NZ, 0.0);
#pragma omp parallel for schedule(static)
        for (int z = 1; z <= NZ - 2; ++z) {
            double t = 0;
            for (int y = 1; y <= NY - 2; ++y)
                for (int x = 1; x <= NX - 2; ++x) {
                    const size_t id = idx(x, y, z);
                    const double a = ux[id], b = uy[id], d = uz[id];
                    t += 0.5 * (a * a + b * b + d * d);
                }
            part[z] = t;
        }
        double t = 0;
        for (int z = 1; z <= NZ - 2; ++z) t += part[z];
        return t;
    }
    void global_sums(double *o) override {
        std::vector<double> part(size_t(NZ) * 5, 0.0);
#pragma omp parallel for schedule(static)
        for (int z = 1; z <= NZ - 2; ++z) {
            double a = 0, b = 0, jx = 0, jy = 0, jz = 0;
            for (int y = 1; y <= NY - 2; ++y)
                for (int x = 1; x <= NX - 2; ++x) {
                    const size_t id = idx(x, y, z);
                    const double r = double(rhor[id]) + double(rhob[id]);
                    a += rhor[id];
                    b += rhob[id];
                    jx += r * ux[id];
                    jy += r * uy[id];
                    jz += r * uz[id];
                }
            double *q = &part[size_t(z) * 5];
            q[0] = a; q[1] = b; q[2] = jx; q[3] = jy; q[4] = jz;
        }
        for (int k = 0; k < 5; ++k) o[k] = 0.0;
        for (int z = 1; z <= NZ - 2; ++z)
            for (int k = 0; k < 5; ++k) o[k] += part[size_t(z) * 5 + k];
    }
    void gradient(int x, int y, int z, double F[3]) override {
        T a, b, d;
        force(x, y, z, a, b, d);
        F[0] = a; F[1] = b; F[2] = d;
    }
    int64_t get_step() override { return step_; }
};

}  // namespace

extern "C" {

void *mcgo_create(const mcg_case *c, int precision) {
    if (!c || c->NX < 4 || c->NY < 4 || c->NZ < 4) return nullptr;
    try {
        if (precision == 32) return static_cast<OracleBase *>(new Oracle<float>(*c));
        return static_cast<OracleBase *>(new Oracle<double>(*c));
    } catch (...) {
        return nullptr;
    }
}
void mcgo_destroy(void *h) { delete static_cast<OracleBase *>(h); }
int mcgo_init_jet(void *h) { return static_cast<OracleBase *>(h)->init(0); }
int mcgo_init_bubble(void *h) { return static_cast<OracleBase *>(h)->init(1); }
int mcgo_init_bubble_tanh(void *h) { return static_cast<OracleBase *>(h)->init(2); }
int mcgo_step(void *h, int n) { return static_cast<OracleBase *>(h)->step(n); }
int64_t mcgo_get_step(void *h) { return static_cast<OracleBase *>(h)->get_step(); }
void mcgo_get_fields(void *h, double *a, double *b, double *x, double *y, double *z) {
    static_cast<OracleBase *>(h)->get_fields(a, b, x, y, z);
}
void mcgo_get_moments(void *h, double *out) { static_cast<OracleBase *>(h)->get_moments(out); }
void mcgo_get_populations(void *h, double *a, double *b) { static_cast<OracleBase *>(h)->get_populations(a, b); }
void mcgo_set_populations(void *h, const double *a, const double *b) { static_cast<OracleBase *>(h)->set_populations(a, b); }
void mcgo_set_densities(void *h, const double *a, const double *b) { static_cast<OracleBase *>(h)->set_densities(a, b); }
void mcgo_get_radial(void *h, double *a, double *b, double *c, double *d, double *e, uint64_t *n, double *p, double *p2) {
    static_cast<OracleBase *>(h)->get_radial(a, b, c, d, e, n, p, p2);
}
void mcgo_stats_info(void *h, int32_t *st, int64_t *ss, int64_t *ns) { static_cast<OracleBase *>(h)->stats_info(st, ss, ns); }
double mcgo_tke(void *h) { return static_cast<OracleBase *>(h)->tke(); }
void mcgo_global_sums(void *h, double *o) { static_cast<OracleBase *>(h)->global_sums(o); }
int mcgo_writeset(void *h, uint8_t *cr, uint8_t *cb) { return static_cast<OracleBase *>(h)->writeset(cr, cb); }
void mcgo_gradient(void *h, int x, int y, int z, double *F) { static_cast<OracleBase *>(h)->gradient(x, y, z, F); }
void mcgo_set_inlet_offset(void *h, const double *du) { static_cast<OracleBase *>(h)->set_inlet_offset(du); }
double mcgo_uy_dot(void *h, const double *w) { return static_cast<OracleBase *>(h)->uy_dot(w); }
void mcgo_branch_masks(void *h, uint8_t *out) { static_cast<OracleBase *>(h)->branch_masks(out); }

// ---- single-node operators, exposed for the invariant tests (FP64) ----
static const Lattice<double> &lat() {
    static const Lattice<double> L;
    return L;
}
void mcgo_lattice(int *cx, int *cy, int *cz, double *w, double *B, double *invc) {
    const auto &L = lat();
    for (int i = 0; i < Q; ++i) {
        cx[i] = CX[i]; cy[i] = CY[i]; cz[i] = CZ[i];
        w[i] = L.w[i]; B[i] = L.B[i]; invc[i] = L.invc[i];
    }
}
double mcgo_feq(int i, double rho, double ux, double uy, double uz) { return feq<double>(lat(), i, rho, ux, uy, uz); }
double mcgo_fneqr(int i, const double *P, double ux, double uy, double uz) { return fneqr<double>(lat(), i, P, ux, uy, uz); }
double mcgo_omega2(int i, double Fx, double Fy, double Fz, double absF, double A) {
    return omega2<double>(lat(), i, Fx, Fy, Fz, absF, A);
}
double mcgo_recolor(int i, double rr, double rb, double Fx, double Fy, double Fz, double absF, double beta) {
    return recolor<double>(lat(), i, rr, rb, Fx, Fy, Fz, absF, beta);
}
int mcgo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
