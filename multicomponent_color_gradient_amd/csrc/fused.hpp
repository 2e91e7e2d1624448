// fused.hpp -- the one-launch step kernel.
//
// Work decomposition: a workgroup owns a tile of 64 x-columns by TY y-rows
// (one wavefront per row, plus three halo waves) and marches along z through
// a segment of planes.  Iteration z
//   A. issues the loads of the lead plane z+2 (pull streaming: 27 or 54 words
//      per node) before anything else, so they are in flight during B;
//   B. collides plane z from its moments (an LDS ring, or registers on
//      statistics steps) and the psi of planes z-1, z, z+1 (a 4-plane LDS
//      ring), and stores the post-collision populations at the node
//      (aligned, coalesced, non-temporal rows);
//   C. reduces the lead plane's words to moments and psi, writes them to the
//      rings, and passes the one barrier of the iteration.  The halo waves
//      gather the one-node ring around the tile (rows y0-1, y0+TY, columns
//      x0-1, x0+64), psi only.
// HBM traffic per interior node is one read and one write of its
// populations; the halo and edge-line re-reads mostly hit L2 because
// neighbouring tiles of a segment run on the same XCD (workgroup ids are
// remapped so that each XCD owns a contiguous block of tiles).  Radial
// statistics, when due, accumulate in an LDS histogram (rows of the tile x
// radial bins) flushed once per workgroup.  The arithmetic per node is the
// same code as the two-pass path, so both paths produce identical
// populations.
#pragma once

#include <algorithm>
#include <vector>

#include "kernels.hpp"
#include "schedule.hpp"

namespace mcg {

// A copy of the run-time strides that the compiler cannot prove loop
// invariant: without it every per-direction address offset is hoisted out of
// the z loop into its own scalar register (~300 of them), which spills.
__device__ __forceinline__ DevParams fresh_strides(const DevParams &P) {
    DevParams Q2 = P;
#define MCG_LAUNDER(f)                                  \
    {                                                   \
        uint32_t v_ = uint32_t(Q2.f);                   \
        asm volatile("" : "+s"(v_));                    \
        Q2.f = decltype(Q2.f)(v_);                      \
    }
    MCG_LAUNDER(NX) MCG_LAUNDER(NY) MCG_LAUNDER(NZ) MCG_LAUNDER(PX) MCG_LAUNDER(IY) MCG_LAUNDER(IZ)
    MCG_LAUNDER(plane) MCG_LAUNDER(rowstride) MCG_LAUNDER(planeel) MCG_LAUNDER(y_mode) MCG_LAUNDER(wrap_gradient)
#undef MCG_LAUNDER
    return Q2;
}

// Pull source of an interior node with the periodic x/y wraps folded into
// per-lane offset corrections within the z-plane and the z wrap into the
// choice of plane (same addresses as source_of()).
struct FastSrc {
    uint32_t p;                   // offset of the node's column within a plane (pinplane)
    uint32_t wxp, wxm, wyp, wym;  // modulo-2^32 offset corrections
    int z, zp, zm;                // the node's plane and the planes it pulls c_z = +1 / -1 populations from
    int o;
    bool yin, yout;
    bool bxp, bxm, bzp, bzm;  // x/z walls: sources across x=0, x=NX-1, z=0, z=NZ-1 bounce back
};

__device__ __forceinline__ FastSrc make_src(const DevParams &P, int x, int y, int z) {
    FastSrc s;
    s.p = pinplane(P, x, y);
    const uint32_t dy = P.rowstride;
    const bool wall = P.wall != 0;
    s.wxp = (!wall && x == 1) ? uint32_t(P.NX - 2) : 0u;
    s.wxm = (!wall && x == P.NX - 2) ? 0u - uint32_t(P.NX - 2) : 0u;
    // with walls, a source across z = 0 or NZ-1 bounces back and its plane is unused
    s.z = z;
    s.zp = (!wall && z == 1) ? P.NZ - 2 : z - 1;
    s.zm = (!wall && z == P.NZ - 2) ? 1 : z + 1;
    s.bxp = wall && x == 1;
    s.bxm = wall && x == P.NX - 2;
    s.bzp = wall && z == 1;
    s.bzm = wall && z == P.NZ - 2;
    const bool jet = P.y_mode == MCG_Y_JET;
    s.wyp = (!jet && y == 1) ? uint32_t(P.NY - 2) * dy : 0u;
    s.wym = (!jet && y == P.NY - 2) ? 0u - uint32_t(P.NY - 2) * dy : 0u;
    s.yin = jet && y == 1;
    s.yout = jet && y == P.NY - 2;
    s.o = z * P.NX + x;
    return s;
}

template <int ST, int I>
__device__ __forceinline__ void fast_load(const DevParams &P, const void *__restrict__ f, const float *__restrict__ bin,
                                          const float *__restrict__ bout, const FastSrc &s, float &fr, float &fb,
                                          float &g) {
    constexpr int cx = kCX[I], cy = kCY[I], cz = kCZ[I];
    const size_t nxz = size_t(P.NX) * P.NZ;
    if constexpr (cy > 0) {
        if (s.yin) {
            const float *b = bin + size_t(2 * slot_of(I)) * nxz;
            const int o = s.o - cz * P.NX - cx;
            fr = b[o];
            fb = b[nxz + o];
            g = fr + fb;
            return;
        }
    }
    if constexpr (cy < 0) {
        if (s.yout) {
            const float *b = bout + size_t(2 * slot_of(I)) * nxz;
            const int o = s.o - cz * P.NX - cx;
            fr = b[o];
            fb = b[nxz + o];
            g = fr + fb;
            return;
        }
    }
    if constexpr (cx != 0 || cz != 0) {
        bool bb = false;
        if constexpr (cx > 0) bb = bb || s.bxp;
        if constexpr (cx < 0) bb = bb || s.bxm;
        if constexpr (cz > 0) bb = bb || s.bzp;
        if constexpr (cz < 0) bb = bb || s.bzm;
        if (bb) {  // half-way bounce-back
            Store<ST>::template load<opp(I)>(plane_ptr(P, f, s.z), uint32_t(P.PX), s.p, fr, fb, g);
            return;
        }
    }
    // unsigned arithmetic: intermediate values may wrap, the final offset is in range
    uint32_t off = s.p - uint32_t(cx) - uint32_t(cy) * P.rowstride;
    if constexpr (cx > 0) off += s.wxp;
    if constexpr (cx < 0) off += s.wxm;
    if constexpr (cy > 0) off += s.wyp;
    if constexpr (cy < 0) off += s.wym;
    const int zs = cz > 0 ? s.zp : (cz < 0 ? s.zm : s.z);
    Store<ST>::template load<I>(plane_ptr(P, f, zs), uint32_t(P.PX), off, fr, fb, g);
}

// Moments of a node of the jet's inlet row (SIDE = +1) or outlet row (SIDE =
// -1): the directions that stream in across that boundary come from its
// buffer (FP32 pairs), the others from the population rows, with an x/z wall
// source taken as the node's own word of the opposite direction (the same
// addresses and values as fast_load).  The source of each direction is fixed
// at compile time, so the code is straight-line and all the loads are in
// flight together; fast_load's per-direction branches waited for the loads of
// each branch in turn (about 17 us per plane at 256x512x256, where a plane of
// the other rows takes 9.5), and the 16-bit step kernels spill 10 to 20 fewer
// VGPRs with it (the jet step 4.5% faster).  FP32: 2% faster when the FP32
// step issues it before its split loads (k_fused), 2% slower after them.
template <int ST, int SIDE>
__device__ __forceinline__ Moments boundary_row_moments(const DevParams &P, const DevBufs &B, const FastSrc &s) {
    const float *bnd = SIDE > 0 ? B.in_src : B.out_src;
    const size_t nxz = size_t(P.NX) * P.NZ;
    MomentAcc acc;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        constexpr int cx = kCX[I], cy = kCY[I], cz = kCZ[I];
        float a, b, g;
        if constexpr (cy == SIDE) {
            const float *bp = bnd + size_t(2 * slot_of(I)) * nxz;
            const int o = s.o - cz * P.NX - cx;
            a = bp[o];
            b = bp[nxz + o];
            g = a + b;
        } else {
            uint32_t off = s.p - uint32_t(cx) - uint32_t(cy) * P.rowstride;
            if constexpr (cx > 0) off += s.wxp;
            if constexpr (cx < 0) off += s.wxm;
            if constexpr (cy > 0) off += s.wyp;
            if constexpr (cy < 0) off += s.wym;
            int zs = cz > 0 ? s.zp : (cz < 0 ? s.zm : s.z);
            if constexpr (cx != 0 || cz != 0) {
                bool bb = false;
                if constexpr (cx > 0) bb = bb || s.bxp;
                if constexpr (cx < 0) bb = bb || s.bxm;
                if constexpr (cz > 0) bb = bb || s.bzp;
                if constexpr (cz < 0) bb = bb || s.bzm;
                if (bb) {
                    off = s.p + uint32_t(opp(I) - I) * uint32_t(P.PX);
                    zs = s.z;
                }
            }
            Store<ST>::template load<I>(plane_ptr(P, B.f_src, zs), uint32_t(P.PX), off, a, b, g);
        }
        acc.add<I>(a, b, g);
    });
    return acc.finish(P.ghalf);
}

template <int ST>
__device__ __forceinline__ Moments fast_moments(const DevParams &P0, const DevBufs &B, int x, int y, int z) {
    const DevParams P = fresh_strides(P0);
    const FastSrc s = make_src(P, x, y, z);
    if (s.yin != s.yout) return s.yin ? boundary_row_moments<ST, 1>(P, B, s) : boundary_row_moments<ST, -1>(P, B, s);
    MomentAcc acc;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b, g;
        fast_load<ST, I>(P, B.f_src, B.in_src, B.out_src, s, a, b, g);
        acc.add<I>(a, b, g);
    });
    return acc.finish(P.ghalf);
}

// psi of any grid position as the colour gradient sees it (0 off the grid,
// the ghost value at ghost positions) and, for interior positions (after the
// optional periodic wrap), the moments of the gathered populations.  The
// colour densities are summed in the order of lbm.cuh:106-115 exactly as the
// moments pass sums them, so psi here and psi from stored moments agree bit
// for bit.  has_m tells whether m was written.
template <int ST>
__device__ __forceinline__ float gather_pos(const DevParams &P0, const DevBufs &B, int x, int y, int z, Moments &m,
                                            bool &has_m) {
    const DevParams P = fresh_strides(P0);
    has_m = false;
    if (x < 0 || x >= P.NX || y < 0 || y >= P.NY || z < 0 || z >= P.NZ) return 0.0f;
    int xn, yn, zn;
    grad_neighbor(P, x, y, z, xn, yn, zn);
    if (!interior(P, xn, yn, zn)) return ghost_psi(P, B, xn, yn, zn);
    m = fast_moments<ST>(P, B, xn, yn, zn);
    has_m = true;
    return psi(m.rr, m.rb);
}

// Split gather for software pipelining: raw_issue starts the 27 or 54 loads
// of an interior node whose sources are all population rows (no inlet or
// outlet source buffer; a wall source is the node's own opposite row),
// raw_reduce turns the words into moments later.  The
// decode and the summation order are those of fast_moments, so the moments
// are bitwise the same.
template <int ST>
struct RawGather {
    uint32_t w[Q * Store<ST>::kWords];
};

template <int ST>
__device__ __forceinline__ void raw_issue(const DevParams &P0, const DevBufs &B, int x, int y, int z, RawGather<ST> &r) {
    const DevParams P = fresh_strides(P0);
    const FastSrc s = make_src(P, x, y, z);
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        constexpr int cx = kCX[I], cy = kCY[I], cz = kCZ[I];
        uint32_t off = s.p - uint32_t(cx) - uint32_t(cy) * P.rowstride;
        if constexpr (cx > 0) off += s.wxp;
        if constexpr (cx < 0) off += s.wxm;
        if constexpr (cy > 0) off += s.wyp;
        if constexpr (cy < 0) off += s.wym;
        int zs = cz > 0 ? s.zp : (cz < 0 ? s.zm : s.z);
        if constexpr (cx != 0 || cz != 0) {
            // half-way bounce-back at an x/z wall: the node's own word of the
            // opposite direction (row opp(I), one PX away), as fast_load reads it
            bool bb = false;
            if constexpr (cx > 0) bb = bb || s.bxp;
            if constexpr (cx < 0) bb = bb || s.bxm;
            if constexpr (cz > 0) bb = bb || s.bzp;
            if constexpr (cz < 0) bb = bb || s.bzm;
            if (bb) {
                off = s.p + uint32_t(opp(I) - I) * uint32_t(P.PX);
                zs = s.z;
            }
        }
        Store<ST>::template load_raw<I>(plane_ptr(P, B.f_src, zs), uint32_t(P.PX), off, r.w + I * Store<ST>::kWords);
    });
}

template <int ST>
__device__ __forceinline__ Moments raw_reduce(const RawGather<ST> &r, float ghalf) {
    MomentAcc acc;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b, g;
        Store<ST>::template decode_raw<I>(r.w + I * Store<ST>::kWords, a, b, g);
        acc.add<I>(a, b, g);
    });
    return acc.finish(ghalf);
}

// psi of a split gather from the colour sums alone, summed in MomentAcc's
// order so that it is bitwise the psi of raw_reduce: all the halo lanes need
// (a third of raw_reduce's arithmetic).
template <int ST>
__device__ __forceinline__ float raw_psi(const RawGather<ST> &r) {
    float sr = 0.f, sb = 0.f;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b, g;
        Store<ST>::template decode_raw<I>(r.w + I * Store<ST>::kWords, a, b, g);
        sr += a;
        sb += b;
    });
    return psi(sr, sb);
}

// Launch shape of the step kernel.  units == nullptr: the uniform shape, nzs
// segments of zlen planes per tile, workgroup b on tile xcd_tile(b).
// Otherwise a banded schedule: workgroup b runs units[b] (nzs then counts its
// z-bands, and instances re-tiled to other rows use nzs uniform segments).
struct FusedGrid {
    int ty;  // core rows per tile of the instance the grid was made for
    int nxt, nyb, nzs, zlen, nwg;
    const FusedUnit *units = nullptr;
    int deal = 0;  // banded: how the tile-planes were dealt to the XCDs (BandDeal)
};

// Tile shape of a k_fused instance: TY core waves (one y-row of 64 x-columns
// each) followed by 3 halo waves, and where the moments of the planes in
// flight live.  The hardware deals the waves of a workgroup to the 4 SIMDs of
// a CU in turn (wave w on SIMD w mod 4, read from the HW_ID register on
// MI355X), so TY a multiple of 4 gives every SIMD the same number of core
// waves; with 9 core rows one SIMD ran 3 of them and the 16-bit step, whose
// arithmetic nearly fills the SIMDs, was 7% slower at 256x512x256.
//   16-bit storage, plain steps: 12 core rows (15 waves, 4 per SIMD, at most
//     128 VGPRs): the halo ring is 156 positions for 768 nodes instead of 148
//     for 512, and every SIMD runs 3 core waves.
//   FP32 (54 words in flight per lane) and the statistics steps: 8 core rows
//     (11 waves, 3 per SIMD).
// kMomLds: the moments of planes z and z+1 wait in an LDS ring instead of
// registers (22 VGPRs fewer, no register rotation per plane).  A statistics
// step keeps its histogram in LDS as well: with RING the ring (68 KB for 8
// rows) and a histogram of up to about 80 KB share the 160 KB; a wider
// histogram takes the instance without RING, moments in registers.  With
// statistics on every step at 256x512x256 the ring ran the FP32 step 8.8%
// and the FP16 step 5.9% faster.
template <int ST, bool STATS, bool RING = true>
struct FusedShape {
    static constexpr bool k16 = ST != MCG_STORE_FP32;
    static constexpr int TY = (k16 && !STATS) ? 12 : 8;
    static constexpr int kWaves = TY + 3;
    static constexpr int kWavesPerEU = (k16 && !STATS) ? 4 : 3;
    static constexpr bool kMomLds = !STATS || RING;
};

// Tile of workgroup b when the hardware deals workgroups round-robin over the
// 8 XCDs: XCD k gets a contiguous range of tiles, so tiles that share halo
// rows and 128-byte edge lines run at the same time behind the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n >> 3, r = n & 7, k = b & 7;
    return k * q + min(k, r) + (b >> 3);
}

// Every lane owns one gather position of the (TY+2) x 66 psi tile: core lanes
// their node, the first halo wave the row y0-1, the second the row y0+TY, the
// third the columns x0-1 and x0+64 (lanes < 2 (TY+2)).  Iteration zi
// gathers the lead plane zi+2 into the psi ring (core lanes keep its moments),
// passes one barrier, and collides plane zi; the first three iterations only
// fill the ring and the moments of planes zb and zb+1.  The gather and the
// collision each appear once in the code, which keeps the loop inside the
// instruction cache.
// MOM: the core lanes also store the moments of their node (B.mom_out), the
// variant the adjoint's recomputation runs.  EXT: x/z walls or the body force
// are on; without them the kernel is compiled with both known to be off.
template <bool STATS, int ST, bool MOM = false, bool EXT = false, bool RING = true>
__global__ __launch_bounds__((64 * FusedShape<ST, STATS, RING>::kWaves))
__attribute__((amdgpu_waves_per_eu(FusedShape<ST, STATS, RING>::kWavesPerEU)))
void k_fused(DevParams P, DevBufs B, FusedGrid G) {
    if constexpr (!EXT) {
        P.wall = 0;
        P.grav = 0.0f;
        P.ghalf = 0.0f;
    }
    using Shape = FusedShape<ST, STATS, RING>;
    constexpr int TY = Shape::TY;
    constexpr bool MLDS = Shape::kMomLds;
    constexpr bool kEarlySerial = MLDS && ST == MCG_STORE_FP32;
    constexpr int H = TY + 2;
    constexpr int W = 66;
    constexpr int NT = 64 * Shape::kWaves;
    constexpr int NMOM = 11;
    __shared__ float phs[4][H][W];
    // moment ring: [plane mod 3][row][moment][x], one 256-byte row per moment
    __shared__ float mring[MLDS ? 3 * TY * NMOM * 64 : 1];
    extern __shared__ unsigned long long hist[];  // STATS: [TY][NR][8], 7 fixed-point sums and the count

    int tile, zb, ze;
    if (G.units) {
        const FusedUnit u = G.units[blockIdx.x];
        tile = int(u.tile);
        zb = int(u.zb);
        ze = int(u.ze);
        if (ze < zb) return;  // padding: the whole workgroup leaves before any barrier
    } else {
        const int t = xcd_tile(int(blockIdx.x), G.nwg);
        tile = t % (G.nxt * G.nyb);
        zb = 1 + (t / (G.nxt * G.nyb)) * G.zlen;
        ze = min(P.NZ - 2, zb + G.zlen - 1);
    }
    const int xt = tile % G.nxt, yb = tile / G.nxt;
    const int tx = int(threadIdx.x), wv = int(threadIdx.y);
    const int lid = wv * 64 + tx;
    const int x0 = 1 + xt * 64, y0 = 1 + yb * TY;
    const bool core = wv < TY;
    const int row = core ? wv : 0;
    const int hrole = core ? -1 : wv - TY;   // halo role 0..2
    const int x = x0 + tx, y = y0 + row;
    const bool node = core && x <= P.NX - 2 && y <= P.NY - 2;

    // this lane's gather position in tile coordinates
    int lr = row + 1, lc = tx + 1;
    bool active = true;
    if (hrole == 0) {
        lr = 0;
    } else if (hrole == 1) {
        lr = H - 1;
    } else if (hrole >= 2) {
        active = tx < 2 * H;
        lr = tx >> 1;
        lc = (tx & 1) ? W - 1 : 0;
    }
    // the tile row whose moments this lane's gather yields (-1: psi only)
    int mrow = core ? row : -1;
    if constexpr (MLDS) {
        // Jet: the inlet row (y = 1) and the outlet row (y = NY-2) gather
        // serially (their sources include the boundary buffers), and with
        // 16-bit storage after the collision, while the workgroup waits at
        // the barrier: the tiles holding them ran 1.8 times as long as the
        // others.  The row-halo wave beside such a row has almost nothing to
        // gather (the ghost inlet row, or a row past the outlet), so the two
        // waves trade positions: the halo wave gathers the boundary row and
        // hands its moments over through the ring, and the core wave only
        // collides.
        if (P.y_mode == MCG_Y_JET) {
            const int rin = 1 - y0, rout = P.NY - 2 - y0;
            if (hrole == 0 && rin >= 0 && rin < TY) {
                lr = rin + 1;
                mrow = rin;
            } else if (hrole == 1 && rout >= 0 && rout < TY && rout != rin) {
                lr = rout + 1;
                mrow = rout;
            } else if (core && row == rin) {
                lr = 0;
                mrow = -1;
            } else if (core && row == rout) {
                lr = H - 1;
                mrow = -1;
            }
        }
    }
    const int gx = x0 - 1 + lc, gy = y0 - 1 + lr;

    if constexpr (STATS) {
        for (int k = lid; k < TY * P.NR * 8; k += NT) hist[k] = 0ull;
    }

    // the lane's lead position after the optional periodic wrap, and whether
    // its gather can be split around the collision (interior, no inlet or
    // outlet source buffer among its sources)
    int gxn = gx, gyn = gy, gzn_unused = 0;
    bool on_grid = gx >= 0 && gx < P.NX && gy >= 0 && gy < P.NY;
    if (on_grid) grad_neighbor(P, gx, gy, 1, gxn, gyn, gzn_unused);
    const bool split_xy = active && on_grid && gxn >= 1 && gxn <= P.NX - 2 && gyn >= 1 && gyn <= P.NY - 2 &&
                          !(P.y_mode == MCG_Y_JET && (gyn == 1 || gyn == P.NY - 2));

    // the moments ring slot of plane q: mring[((q mod 3) * TY + row) * NMOM + k][tx]
    auto mslot = [&](int q) -> float * { return mring + (size_t((q % 3) * TY + row) * NMOM) * 64 + tx; };

    Moments cur{}, nx1{}, nx2{};
    RawGather<ST> raw;
    // iteration zi collides plane zi (once zi >= zb) and gathers the lead plane zi+2;
    // the three fill iterations gather planes zb-1, zb and zb+1
    for (int zi = zb - 3; zi <= ze; ++zi) {
        // the plane index passes through an opaque copy so that addresses are
        // recomputed per plane instead of becoming loop-carried pointers
        int z = zi;
        asm volatile("" : "+s"(z));
        // marked uniform again, so that the plane bases stay scalar
        z = __builtin_amdgcn_readfirstlane(z);
        const int zl = z + 2;
        const bool lead = zl <= ze + 1;
        int zln = zl;
        if (P.wrap_gradient && !P.wall) zln = wrapN(zl, P.NZ);
        const bool split = lead && split_xy && zln >= 1 && zln <= P.NZ - 2;
        // FP32: a lane that cannot split its gather (positions off the
        // interior, the jet's boundary rows) gathers before the collision, so
        // its wave waits on the loads while the other waves of its SIMD
        // collide (at 256x512x256 the jet step ran 3% faster, the rising
        // bubble 4%).  Without walls it also goes before the split loads are
        // issued, so their words are not live during it (the jet 2% faster
        // again); with walls after them (the rising bubble ran 2.6% slower
        // the other way).  With 16-bit storage the early gather spills 12 to
        // 40 more VGPRs.  The ring slots of plane zl are read by no one before
        // the barrier.
        // (the same block twice, one compiled per instance: as a shared lambda
        // it ran both orders slower)
        if constexpr (kEarlySerial && !EXT) {
            if (lead && active && !split) {
                bool has_m;
                const float ph = gather_pos<ST>(P, B, gx, gy, zl, nx2, has_m);
                if (mrow >= 0) {
                    float *ms = mring + (size_t((zl % 3) * TY + mrow) * NMOM) * 64 + tx;
                    const float v[NMOM] = {nx2.rr, nx2.rb, nx2.ux, nx2.uy, nx2.uz, nx2.pxx,
                                           nx2.pxy, nx2.pyy, nx2.pyz, nx2.pzz, nx2.pxz};
                    for (int k = 0; k < NMOM; ++k) ms[k * 64] = v[k];
                }
                phs[zl & 3][lr][lc] = ph;
            }
        }
        if (split) raw_issue<ST>(P, B, gxn, gyn, zln, raw);
        if constexpr (kEarlySerial && EXT) {
            if (lead && active && !split) {
                bool has_m;
                const float ph = gather_pos<ST>(P, B, gx, gy, zl, nx2, has_m);
                if (mrow >= 0) {
                    float *ms = mring + (size_t((zl % 3) * TY + mrow) * NMOM) * 64 + tx;
                    const float v[NMOM] = {nx2.rr, nx2.rb, nx2.ux, nx2.uy, nx2.uz, nx2.pxx,
                                           nx2.pxy, nx2.pyy, nx2.pyz, nx2.pzz, nx2.pxz};
                    for (int k = 0; k < NMOM; ++k) ms[k * 64] = v[k];
                }
                phs[zl & 3][lr][lc] = ph;
            }
        }
        if (node && z >= zb) {
            if constexpr (MLDS) {
                const float *ms = mslot(z);
                cur = Moments{ms[0 * 64], ms[1 * 64], ms[2 * 64], ms[3 * 64], ms[4 * 64], ms[5 * 64],
                              ms[6 * 64], ms[7 * 64], ms[8 * 64], ms[9 * 64], ms[10 * 64]};
            }
            const Moments &m = cur;
            auto phi = [&](auto Ic) -> float {
                constexpr int I = decltype(Ic)::value;
                return phs[(z + kCZ[I]) & 3][row + 1 + kCY[I]][tx + 1 + kCX[I]];
            };
            const NodeColl n = node_setup(P, m, y, phi);
            const DevParams PL = fresh_strides(P);
            collide_store<ST>(n, plane_ptr(PL, B.f_dst, z), uint32_t(PL.PX), pinplane(PL, x, y), pidx(PL, x, y, z),
                              sr_step_key(B.step));
            if constexpr (MOM) {
                const uint32_t q = midx(PL, x, y, z);
                const size_t pl = PL.plane;
                const float v[11] = {m.rr, m.rb, m.ux, m.uy, m.uz, m.pxx, m.pxy, m.pyy, m.pyz, m.pzz, m.pxz};
                for (int k = 0; k < 11; ++k) __builtin_nontemporal_store(v[k], &B.mom_out[k * pl + q]);
            }
            if (P.y_mode == MCG_Y_JET) {
                if (y == 1 && is_jet(P, x, z)) inlet_emit(P, B.in_dst, m, x, z, B.step, B.inlet_du);
                if (y == P.NY - 2) outlet_emit(P, B.gB_src, B.gB_dst, B.out_dst, m, x, z);
            }
            if constexpr (STATS) {
                float dx, dz, rr;
                const int bin = radial_bin(P, x, z, dx, dz, rr);
                if (bin >= 0) {
                    float ur = 0.0f;
                    if (rr > 1.0e-12f) ur = (m.ux * dx + m.uz * dz) / rr;
                    stat_sample(hist + (size_t(row) * P.NR + bin) * 8, 1, double(m.uy), double(ur),
                                double(psi(m.rr, m.rb)));
                }
            }
        }
        if (lead && active && (!kEarlySerial || split)) {
            float ph = 0.0f;
            if (split && !core) {
                ph = raw_psi<ST>(raw);  // a halo wave: only psi is used
            } else if (split) {
                nx2 = raw_reduce<ST>(raw, P.ghalf);
                ph = psi(nx2.rr, nx2.rb);
            } else if constexpr (!kEarlySerial) {
                bool has_m;
                ph = gather_pos<ST>(P, B, gx, gy, zl, nx2, has_m);
            }
            if constexpr (MLDS) {
                if (mrow >= 0) {
                    float *ms = mring + (size_t((zl % 3) * TY + mrow) * NMOM) * 64 + tx;
                    const float v[NMOM] = {nx2.rr, nx2.rb, nx2.ux, nx2.uy, nx2.uz, nx2.pxx,
                                           nx2.pxy, nx2.pyy, nx2.pyz, nx2.pzz, nx2.pxz};
                    for (int k = 0; k < NMOM; ++k) ms[k * 64] = v[k];
                }
            }
            phs[zl & 3][lr][lc] = ph;
        }
        __syncthreads();
        if constexpr (!MLDS) {
            cur = nx1;
            nx1 = nx2;
        }
    }

    if constexpr (STATS) {
        __syncthreads();
        const size_t nb = size_t(P.NY) * P.NR;
        for (int k = lid; k < TY * P.NR; k += NT) {
            const int row = k / P.NR, bin = k % P.NR;
            const int yy = y0 + row;
            const unsigned long long *hb = hist + size_t(k) * 8;
            const unsigned long long c = hb[7];
            if (c == 0 || yy > P.NY - 2) continue;
            const size_t o = size_t(yy) * P.NR + bin;
            for (int q = 0; q < 7; ++q) atomicAdd(&B.stats_fx[q * nb + o], hb[q]);
            atomicAdd(&B.count[o], c);
        }
    }
}

// core rows of the plain (no statistics) step of a storage: the shape the
// launch-shape timing and the segment counts refer to
inline int fused_plain_ty(int storage) {
    return storage == MCG_STORE_FP32 ? FusedShape<MCG_STORE_FP32, false>::TY : FusedShape<MCG_STORE_FP16, false>::TY;
}

// The launch of zs z segments: tiles of 64 x-columns by ty rows, each
// marching through a segment of ceil(IZ/zs) planes.
inline FusedGrid fused_grid_for(const DevParams &P, int zs, int ty) {
    FusedGrid G;
    G.ty = ty;
    G.nxt = P.PX / 64;
    G.nyb = (P.IY + ty - 1) / ty;
    G.zlen = (P.IZ + zs - 1) / zs;
    G.nzs = (P.IZ + G.zlen - 1) / G.zlen;
    G.nwg = G.nxt * G.nyb * G.nzs;
    return G;
}

// Segment counts worth trying for the plain step: every distinct segmentation
// with segments of at least 8 planes (at most 64 segments).  Each segment costs
// 2 fill planes of gathers, and the best count depends on which tiles run at
// the same time: at 256x512x256 the 16-bit step (12-row tiles) takes 1.96 ms
// with 15 segments, 2.00 with 7, 9 or 13, 2.10 with 16 and 2.36 with 6, and
// the FP32 step 3.25 ms with 4, 3.30 with 5 and 3.60 with 29, whatever the
// padding of the row blocks (32 to 320 words).  The solver times every
// candidate on its first steps and keeps the fastest; the results do not
// depend on the segmentation.
inline std::vector<FusedGrid> fused_candidates_ty(const DevParams &P, int ty, int max_segments) {
    std::vector<FusedGrid> all;
    const int maxseg = std::max(1, std::min(max_segments, P.IZ / 8));
    for (int zs = 1; zs <= maxseg; ++zs) {
        const FusedGrid G = fused_grid_for(P, zs, ty);
        if (all.empty() || all.back().nzs != G.nzs) all.push_back(G);
    }
    return all;
}
inline std::vector<FusedGrid> fused_candidates(const DevParams &P) {
    return fused_candidates_ty(P, fused_plain_ty(P.storage), 64);
}

// The banded candidates of a launch: the kBandedCandidates band rules of
// shortest estimated span (banded_rules) for the plain instance's tiles.
inline std::vector<BandedSchedule> banded_candidates(const DevParams &P, int num_cu) {
    const int ty = fused_plain_ty(P.storage);
    const int ntiles = (P.PX / 64) * ((P.IY + ty - 1) / ty);
    const double fill = P.storage == MCG_STORE_FP32 ? 1.5 : 1.9;
    return banded_rules(ntiles, P.IZ, std::max(1, num_cu / kXcds), fill);
}

// The segmentation used before (or without) the timing: among the counts that
// give at least kMinRounds rounds of resident workgroups (4 for FP32, 16 for
// 16-bit storage; one workgroup per CU), the fewest rounds x (segment length +
// 2 fill planes); grids too small for that many rounds take the most segments.
inline FusedGrid fused_grid(const DevParams &P, int num_cu) {
    const int min_rounds = P.storage == MCG_STORE_FP32 ? 4 : 16;
    const int slots = std::max(1, num_cu);
    FusedGrid best{}, most{};
    double best_cost = 1e300;
    for (const FusedGrid &G : fused_candidates(P)) {
        most = G;
        const int rounds = (G.nwg + slots - 1) / slots;
        if (rounds < min_rounds) continue;
        const double cost = double(rounds) * double(G.zlen + 2);
        if (cost < best_cost) {
            best_cost = cost;
            best = G;
        }
    }
    return best_cost < 1e300 ? best : most;
}

// the statistics instances (8-row tiles) keep their histogram in dynamic LDS
// next to the psi ring and, with RING, the moment ring
constexpr int kStatsTY = FusedShape<MCG_STORE_FP32, true>::TY;
static_assert(FusedShape<MCG_STORE_FP16, true>::TY == kStatsTY, "one statistics tile height");
inline size_t fused_stats_lds(const DevParams &P) { return size_t(kStatsTY) * P.NR * 8 * sizeof(unsigned long long); }
inline size_t fused_static_lds(bool ring) {
    return size_t(4) * (kStatsTY + 2) * 66 * sizeof(float) + (ring ? size_t(3) * kStatsTY * 11 * 64 * sizeof(float) : 0);
}
// the histogram fits in the LDS a workgroup may use: next to the psi ring
// (some statistics instance runs), and next to the moment ring as well
inline bool fused_has_stats(const DevParams &P, size_t lds_per_block) {
    return fused_stats_lds(P) + fused_static_lds(false) <= lds_per_block;
}
inline bool fused_stats_ring(const DevParams &P, size_t lds_per_block) {
    return fused_stats_lds(P) + fused_static_lds(true) <= lds_per_block;
}

// grid G for an instance of ty rows: G itself, or uniform segments of ty-row
// tiles with G's segment (or band) count
inline FusedGrid fused_grid_for_instance(const DevParams &P, const FusedGrid &G, int ty) {
    return G.ty == ty ? G : fused_grid_for(P, G.nzs, ty);
}

// one instance on grid G, re-tiled to the instance's rows if they differ
template <bool STATS, int ST, bool MOM, bool EXT, bool RING = true>
inline void launch_fused_one(const DevParams &P, const DevBufs &B, hipStream_t stream, const FusedGrid &G, size_t lds) {
    using Shape = FusedShape<ST, STATS, RING>;
    const FusedGrid Gi = fused_grid_for_instance(P, G, Shape::TY);
    hipLaunchKernelGGL((k_fused<STATS, ST, MOM, EXT, RING>), dim3(Gi.nwg), dim3(64, Shape::kWaves, 1), lds, stream, P, B,
                       Gi);
}

template <int ST>
inline int launch_fused_st(const DevParams &P, const DevBufs &B, hipStream_t stream, const FusedGrid &G) {
    if (P.wall || P.grav != 0.0f) {
        if (B.do_stats && B.stats_ring) {
            launch_fused_one<true, ST, false, true, true>(P, B, stream, G, fused_stats_lds(P));
        } else if (B.do_stats) {
            launch_fused_one<true, ST, false, true, false>(P, B, stream, G, fused_stats_lds(P));
        } else if (B.mom_out) {
            launch_fused_one<false, ST, true, true>(P, B, stream, G, 0);
        } else {
            launch_fused_one<false, ST, false, true>(P, B, stream, G, 0);
        }
    } else if (B.do_stats && B.stats_ring) {
        launch_fused_one<true, ST, false, false, true>(P, B, stream, G, fused_stats_lds(P));
    } else if (B.do_stats) {
        launch_fused_one<true, ST, false, false, false>(P, B, stream, G, fused_stats_lds(P));
    } else if (B.mom_out) {
        launch_fused_one<false, ST, true, false>(P, B, stream, G, 0);
    } else {
        launch_fused_one<false, ST, false, false>(P, B, stream, G, 0);
    }
    return int(hipGetLastError());
}

// dynamic LDS above the default limit must be allowed per kernel
template <int ST>
inline int fused_allow_lds(size_t bytes) {
    for (const void *k : {reinterpret_cast<const void *>(&k_fused<true, ST, false, false, true>),
                          reinterpret_cast<const void *>(&k_fused<true, ST, false, false, false>),
                          reinterpret_cast<const void *>(&k_fused<true, ST, false, true, true>),
                          reinterpret_cast<const void *>(&k_fused<true, ST, false, true, false>)})
        if (int rc = int(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)))) return rc;
    return 0;
}

inline int launch_fused(const DevParams &P, const DevBufs &B, hipStream_t stream, const FusedGrid &G) {
    switch (P.storage) {
        case MCG_STORE_FP16: return launch_fused_st<MCG_STORE_FP16>(P, B, stream, G);
        case MCG_STORE_BF16: return launch_fused_st<MCG_STORE_BF16>(P, B, stream, G);
        default: return launch_fused_st<MCG_STORE_FP32>(P, B, stream, G);
    }
}

}  // namespace mcg
