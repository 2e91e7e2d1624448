// fused.hpp -- the one-launch step kernel.
//
// Work decomposition: a workgroup owns a tile of 64 x-columns by TY y-rows
// (one wavefront per row) and marches along z through a segment of planes.
// For each plane z it
//   A. gathers the 54 populations of its nodes on the lead plane z+1 (pull
//      streaming), reduces them to moments kept in registers, and writes
//      psi = (rho_r - rho_b)/(rho_r + rho_b) of the lead plane into a 4-deep
//      LDS ring; the one-node halo ring (rows y0-1, y0+TY, columns x0-1,
//      x0+64) is gathered too, densities only;
//   B. after one barrier, collides plane z from the moments held in registers
//      and the psi of planes z-1, z, z+1 read from LDS, and stores the 54
//      post-collision populations at the node (aligned, coalesced rows).
// HBM traffic per interior node is one read and one write of 54 values; the
// halo re-reads hit L2 because neighbouring tiles of a segment run on the
// same XCD (workgroup ids are remapped so that each XCD owns a contiguous
// block of tiles).  Radial statistics, when due, accumulate in an LDS
// histogram (rows of the tile x radial bins) flushed once per workgroup.
// The arithmetic per node is the same code as the two-pass path, so both
// paths produce identical populations.
#pragma once

#include <algorithm>

#include "kernels.hpp"

namespace mcg {

// density-based psi of interior node (x,y,z), summed in the order of
// lbm.cuh:106-115 so that it equals psi of the moments pass bit for bit
template <int ST>
__device__ __forceinline__ float gathered_psi(const DevParams &P, const DevBufs &B, int x, int y, int z) {
    float sr = 0.f, sb = 0.f;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b;
        gather<ST, I>(P, B.f_src, B.in_src, B.out_src, x, y, z, a, b);
        sr += a;
        sb += b;
    });
    return psi(sr, sb);
}

// psi as the colour gradient sees it at any grid position; 0 off the grid
// out of line: called for the halo ring and ghost positions only, it keeps the
// generic gather's address arithmetic out of the main loop's register budget
template <int ST>
__device__ __noinline__ float psi_at(const DevParams &P, const DevBufs &B, int x, int y, int z) {
    if (x < 0 || x >= P.NX || y < 0 || y >= P.NY || z < 0 || z >= P.NZ) return 0.0f;
    int xn, yn, zn;
    grad_neighbor(P, x, y, z, xn, yn, zn);
    if (interior(P, xn, yn, zn)) return gathered_psi<ST>(P, B, xn, yn, zn);
    return ghost_psi(P, B, xn, yn, zn);
}

// Pull source of an interior node with the periodic wraps folded into
// per-lane offset corrections (same addresses as source_of()).
struct FastSrc {
    uint32_t p;
    uint32_t wxp, wxm, wyp, wym, wzp, wzm;  // modulo-2^32 offset corrections
    int o;
    bool yin, yout;
};

__device__ __forceinline__ FastSrc make_src(const DevParams &P, int x, int y, int z) {
    FastSrc s;
    s.p = pidx(P, x, y, z);
    const uint32_t dy = P.rowstride, dz = P.rowstride * uint32_t(P.IY);
    s.wxp = (x == 1) ? uint32_t(P.NX - 2) : 0u;
    s.wxm = (x == P.NX - 2) ? 0u - uint32_t(P.NX - 2) : 0u;
    s.wzp = (z == 1) ? uint32_t(P.NZ - 2) * dz : 0u;
    s.wzm = (z == P.NZ - 2) ? 0u - uint32_t(P.NZ - 2) * dz : 0u;
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
                                          const float *__restrict__ bout, const FastSrc &s, float &fr, float &fb) {
    constexpr int cx = kCX[I], cy = kCY[I], cz = kCZ[I];
    const size_t nxz = size_t(P.NX) * P.NZ;
    if constexpr (cy > 0) {
        if (s.yin) {
            const float *b = bin + size_t(2 * slot_of(I)) * nxz;
            const int o = s.o - cz * P.NX - cx;
            fr = b[o];
            fb = b[nxz + o];
            return;
        }
    }
    if constexpr (cy < 0) {
        if (s.yout) {
            const float *b = bout + size_t(2 * slot_of(I)) * nxz;
            const int o = s.o - cz * P.NX - cx;
            fr = b[o];
            fb = b[nxz + o];
            return;
        }
    }
    // unsigned arithmetic: intermediate values may wrap, the final offset is in range
    uint32_t off = s.p - uint32_t(cx) - uint32_t(cy) * P.rowstride - uint32_t(cz) * (P.rowstride * uint32_t(P.IY));
    if constexpr (cx > 0) off += s.wxp;
    if constexpr (cx < 0) off += s.wxm;
    if constexpr (cy > 0) off += s.wyp;
    if constexpr (cy < 0) off += s.wym;
    if constexpr (cz > 0) off += s.wzp;
    if constexpr (cz < 0) off += s.wzm;
    Store<ST>::template load<I>(f, uint32_t(P.PX), off, fr, fb);
}

template <int ST>
__device__ __forceinline__ Moments fast_moments(const DevParams &P, const DevBufs &B, int x, int y, int z) {
    const FastSrc s = make_src(P, x, y, z);
    MomentAcc acc;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b;
        fast_load<ST, I>(P, B.f_src, B.in_src, B.out_src, s, a, b);
        acc.add<I>(a, b);
    });
    return acc.finish();
}

struct FusedGrid {
    int nxt, nyb, nzs, zlen, nwg;
};

#ifndef MCG_FUSED_WAVES_PER_EU
#define MCG_FUSED_WAVES_PER_EU 1
#endif

template <int TY, bool STATS, int ST>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(MCG_FUSED_WAVES_PER_EU)))
void k_fused(DevParams P, DevBufs B, FusedGrid G) {
    constexpr int H = TY + 2;
    constexpr int W = 66;
    __shared__ float phs[4][H][W];
    extern __shared__ double hist[];  // STATS: [TY][NR][8], the last slot a u64 count

    int t = int(blockIdx.x);
    if ((G.nwg & 7) == 0) t = (t & 7) * (G.nwg >> 3) + (t >> 3);  // contiguous tiles per XCD
    const int xt = t % G.nxt;
    const int r1 = t / G.nxt;
    const int yb = r1 % G.nyb;
    const int zs = r1 / G.nyb;
    const int tx = int(threadIdx.x), ty = int(threadIdx.y);
    const int lid = ty * 64 + tx;
    const int x0 = 1 + xt * 64, y0 = 1 + yb * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x <= P.NX - 2 && y <= P.NY - 2;
    const int zb = 1 + zs * G.zlen;
    const int ze = min(P.NZ - 2, zb + G.zlen - 1);

    if constexpr (STATS) {
        for (int k = lid; k < TY * P.NR * 8; k += 64 * TY) hist[k] = 0.0;
    }

    // halo ring of one plane: rows y0-1 and y0+TY (64 columns), columns x0-1 and x0+64 (H rows)
    constexpr int NH = 128 + 2 * H;
    auto halo = [&](int zz) {
        for (int h = lid; h < NH; h += 64 * TY) {
            int r, c;
            if (h < 64) {
                r = 0;
                c = h + 1;
            } else if (h < 128) {
                r = H - 1;
                c = h - 63;
            } else {
                r = (h - 128) >> 1;
                c = ((h - 128) & 1) ? W - 1 : 0;
            }
            phs[zz & 3][r][c] = psi_at<ST>(P, B, x0 - 1 + c, y0 - 1 + r, zz);
        }
    };

    // prologue: psi of plane zb-1, moments and psi of plane zb
    phs[(zb - 1) & 3][ty + 1][tx + 1] = psi_at<ST>(P, B, x, y, zb - 1);
    halo(zb - 1);
    Moments cur{};
    if (valid) {
        cur = fast_moments<ST>(P, B, x, y, zb);
        phs[zb & 3][ty + 1][tx + 1] = psi(cur.rr, cur.rb);
    } else {
        phs[zb & 3][ty + 1][tx + 1] = psi_at<ST>(P, B, x, y, zb);
    }
    halo(zb);

    for (int z = zb; z <= ze; ++z) {
        // A: lead plane
        const int zl = z + 1;
        Moments nxt{};
        if (valid && zl <= ze) {
            nxt = fast_moments<ST>(P, B, x, y, zl);
            phs[zl & 3][ty + 1][tx + 1] = psi(nxt.rr, nxt.rb);
        } else {
            phs[zl & 3][ty + 1][tx + 1] = psi_at<ST>(P, B, x, y, zl);
        }
        halo(zl);
        __syncthreads();
        // B: collide plane z
        if (valid) {
            auto phi = [&](auto Ic) -> float {
                constexpr int I = decltype(Ic)::value;
                return phs[(z + kCZ[I]) & 3][ty + 1 + kCY[I]][tx + 1 + kCX[I]];
            };
            const NodeColl n = node_setup(P, cur, y, phi);
            const uint32_t p = pidx(P, x, y, z);
            unroll<0, Q>([&](auto Ic) {
                constexpr int I = decltype(Ic)::value;
                float a, b;
                collide_dir<I>(n, a, b);
                Store<ST>::template store<I>(B.f_dst, uint32_t(P.PX), p, a, b);
            });
            if (P.y_mode == MCG_Y_JET) {
                if (y == 1 && is_jet(P, x, z)) inlet_emit(P, B.in_dst, cur, x, z);
                if (y == P.NY - 2) outlet_emit(P, B.gB_src, B.gB_dst, B.out_dst, cur, x, z);
            }
            if constexpr (STATS) {
                float dx, dz, r;
                const int bin = radial_bin(P, x, z, dx, dz, r);
                if (bin >= 0) {
                    float ur = 0.0f;
                    if (r > 1.0e-12f) ur = (cur.ux * dx + cur.uz * dz) / r;
                    const double vy = double(cur.uy), vr = double(ur), ph = double(psi(cur.rr, cur.rb));
                    double *hb = hist + (size_t(ty) * P.NR + bin) * 8;
                    atomicAdd(hb + 0, vy);
                    atomicAdd(hb + 1, vy * vy);
                    atomicAdd(hb + 2, vr);
                    atomicAdd(hb + 3, vr * vr);
                    atomicAdd(hb + 4, vr * vy);
                    atomicAdd(hb + 5, ph);
                    atomicAdd(hb + 6, ph * ph);
                    atomicAdd(reinterpret_cast<unsigned long long *>(hb + 7), 1ull);
                }
            }
        }
        cur = nxt;
    }

    if constexpr (STATS) {
        __syncthreads();
        const size_t nb = size_t(P.NY) * P.NR;
        for (int k = lid; k < TY * P.NR; k += 64 * TY) {
            const int row = k / P.NR, bin = k % P.NR;
            const int yy = y0 + row;
            const double *hb = hist + size_t(k) * 8;
            const unsigned long long c = reinterpret_cast<const unsigned long long *>(hb)[7];
            if (c == 0 || yy > P.NY - 2) continue;
            const size_t o = size_t(yy) * P.NR + bin;
            for (int q = 0; q < 7; ++q) atomicAdd(&B.stats[q * nb + o], hb[q]);
            atomicAdd(&B.count[o], c);
        }
    }
}

// rows per workgroup of the fused kernel
constexpr int kFusedTY = 8;

inline FusedGrid fused_grid(const DevParams &P, int zsegs) {
    FusedGrid G;
    G.nxt = P.PX / 64;
    G.nyb = (P.IY + kFusedTY - 1) / kFusedTY;
    G.nzs = zsegs;
    G.zlen = (P.IZ + zsegs - 1) / zsegs;
    G.nzs = (P.IZ + G.zlen - 1) / G.zlen;
    G.nwg = G.nxt * G.nyb * G.nzs;
    return G;
}

// z segments: aim at ~3 workgroups per CU over 256 CUs, segments of >= 16 planes
inline int fused_zsegs(const DevParams &P) {
    const int tiles = (P.PX / 64) * ((P.IY + kFusedTY - 1) / kFusedTY);
    int zs = (768 + tiles - 1) / tiles;
    zs = std::max(1, std::min(zs, P.IZ / 16));
    return std::max(1, zs);
}

inline bool fused_has_stats(const DevParams &P) { return size_t(kFusedTY) * P.NR * 8 * sizeof(double) <= 65536; }

template <int ST>
inline int launch_fused_st(const DevParams &P, const DevBufs &B, hipStream_t stream) {
    const FusedGrid G = fused_grid(P, fused_zsegs(P));
    const dim3 block(64, kFusedTY, 1);
    if (B.do_stats) {
        const size_t shm = size_t(kFusedTY) * P.NR * 8 * sizeof(double);
        hipLaunchKernelGGL((k_fused<kFusedTY, true, ST>), dim3(G.nwg), block, shm, stream, P, B, G);
    } else {
        hipLaunchKernelGGL((k_fused<kFusedTY, false, ST>), dim3(G.nwg), block, 0, stream, P, B, G);
    }
    return int(hipGetLastError());
}

inline int launch_fused(const DevParams &P, const DevBufs &B, hipStream_t stream) {
    switch (P.storage) {
        case MCG_STORE_FP16: return launch_fused_st<MCG_STORE_FP16>(P, B, stream);
        case MCG_STORE_BF16: return launch_fused_st<MCG_STORE_BF16>(P, B, stream);
        default: return launch_fused_st<MCG_STORE_FP32>(P, B, stream);
    }
}

}  // namespace mcg
