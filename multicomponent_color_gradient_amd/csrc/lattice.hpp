// lattice.hpp -- D3Q27 tables as compile-time constants for gfx950 kernels.
//
// Direction order, weights, perturbation coefficients B_i, Hermite
// polynomials and 1/|c_i| follow stencil_ct.cuh:9-203 / stencil.cuh:8-24 of
// the reference: the population layout of the drop-in (population index
// i*Ncells + id) depends on this order, so it is kept.  All constants are
// evaluated in float, as constants.cuh:45-53 does for real_t=float.
#pragma once

#include <cstdint>
#include <type_traits>

namespace mcg {

constexpr int Q = 27;

constexpr int kCX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int kCY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int kCZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

// the 9 directions with c_y = +1 (fed by the inlet) and c_y = -1 (fed by the
// outlet), in increasing direction order; slot k of a boundary buffer holds
// direction kUp[k] / kDown[k]
constexpr int kUp[9] = {3, 7, 11, 14, 17, 19, 21, 24, 25};
constexpr int kDown[9] = {4, 8, 12, 13, 18, 20, 22, 23, 26};

constexpr int slot_of(int i) {
    for (int k = 0; k < 9; ++k)
        if (kUp[k] == i || kDown[k] == i) return k;
    return -1;
}

constexpr float cs2 = float(1.0 / 3.0);
constexpr float cs4 = cs2 * cs2;
constexpr float cs6 = cs4 * cs2;
constexpr float inv_cs2 = 1.0f / cs2;
constexpr float inv_2cs2 = 1.0f / (2.0f * cs2);
constexpr float inv_2cs4 = 1.0f / (2.0f * cs4);
constexpr float inv_6cs6 = 1.0f / (6.0f * cs6);
constexpr float inv_2cs6 = 1.0f / (2.0f * cs6);

constexpr int csq(int i) { return kCX[i] * kCX[i] + kCY[i] * kCY[i] + kCZ[i] * kCZ[i]; }
// opposite direction: the directions 2k-1 and 2k (k = 1..13) are opposite
constexpr int opp(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }
static_assert(kCX[opp(25)] == -kCX[25] && kCY[opp(13)] == -kCY[13] && kCZ[opp(22)] == -kCZ[22], "opposites");

constexpr float w(int i) {
    return csq(i) == 0 ? float(8.0 / 27.0) : csq(i) == 1 ? float(2.0 / 27.0) : csq(i) == 2 ? float(1.0 / 54.0) : float(1.0 / 216.0);
}
constexpr float Bc(int i) { return i == 0 ? float(-10.0 / 27.0) : w(i); }
constexpr float invc(int i) {
    return csq(i) == 0 ? 0.0f : csq(i) == 1 ? 1.0f : csq(i) == 2 ? float(0.7071067811865475) : float(0.5773502691896258);
}

constexpr float cxf(int i) { return float(kCX[i]); }
constexpr float cyf(int i) { return float(kCY[i]); }
constexpr float czf(int i) { return float(kCZ[i]); }

// second-order Hermite polynomials
constexpr float Hxx(int i) { return float(kCX[i] * kCX[i]) - cs2; }
constexpr float Hxy(int i) { return float(kCX[i] * kCY[i]); }
constexpr float Hyy(int i) { return float(kCY[i] * kCY[i]) - cs2; }
constexpr float Hyz(int i) { return float(kCY[i] * kCZ[i]); }
constexpr float Hzz(int i) { return float(kCZ[i] * kCZ[i]) - cs2; }
constexpr float Hxz(int i) { return float(kCX[i] * kCZ[i]); }
// third-order Hermite polynomials
constexpr float Hxxy(int i) { return float(kCX[i] * kCX[i] * kCY[i]) - cs2 * float(kCY[i]); }
constexpr float Hxxz(int i) { return float(kCX[i] * kCX[i] * kCZ[i]) - cs2 * float(kCZ[i]); }
constexpr float Hxyy(int i) { return float(kCX[i] * kCY[i] * kCY[i]) - cs2 * float(kCX[i]); }
constexpr float Hxzz(int i) { return float(kCX[i] * kCZ[i] * kCZ[i]) - cs2 * float(kCX[i]); }
constexpr float Hyyz(int i) { return float(kCY[i] * kCY[i] * kCZ[i]) - cs2 * float(kCZ[i]); }
constexpr float Hyzz(int i) { return float(kCY[i] * kCZ[i] * kCZ[i]) - cs2 * float(kCY[i]); }
constexpr float Hxyz(int i) { return float(kCX[i] * kCY[i] * kCZ[i]); }

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__host__ __device__ __forceinline__ void unroll(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        unroll<B + 1, E>(f);
    }
}

}  // namespace mcg
