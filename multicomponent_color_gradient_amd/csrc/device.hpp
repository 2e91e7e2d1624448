// device.hpp -- device-side state description, HBM layout and the per-node
// streaming (pull) access used by every step kernel.
//
// HBM layout (FP32 storage)
//   populations  : two buffers (read "src", write "dst"), each 54 planes
//                  [colour*27 + i][plane], plane = PX * (NY-2) * (NZ-2),
//                  interior nodes only, x fastest with the row padded to PX
//                  (a multiple of 64) so that every wavefront row is aligned.
//                  A buffer holds the post-collision populations of the last
//                  step stored at their SOURCE node: the step gathers
//                  f_i(x) = f*_i(x - c_i) (pull), which reproduces the
//                  reference's in-place push (lbm.cuh:230-244) exactly.
//   inlet/outlet : [slot k][colour][NZ][NX], k indexes the 9 directions with
//                  c_y=+1 (inlet) / c_y=-1 (outlet).  Entry (x,z) holds the
//                  population that the ghost node (x,0,z) / (x,NY-1,z) sends
//                  into the first/last fluid plane, double buffered.  Entries
//                  the reference never writes keep their initial values
//                  (quirks Q1/Q2: jet/boundary_conditions.cuh:38-41 and the
//                  unwrapped targets of :73,:222).
//   outlet ghost : [5][NZ][NX] rho_r, rho_b, u_x, u_y, u_z of plane y=NY-1,
//                  double buffered (the colour gradient of step n reads the
//                  values of step n-1, jet/boundary_conditions.cuh:186-190).
//   ghost rho    : full-grid rho_r, rho_b holding the values of ghost nodes
//                  (x/z frames and the y=0 plane), read by the colour
//                  gradient without wrap (forces.cuh:31-33).
//   moments      : 11 interior planes (rho_r, rho_b, u, Pi), written by the
//                  two-pass path and by the on-demand moments pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mcg.h"
#include "physics.hpp"

namespace mcg {

struct DevParams {
    int NX, NY, NZ;
    int PX;           // padded interior row pitch (elements)
    int IY, IZ;       // NY-2, NZ-2
    uint32_t plane;   // PX*IY*IZ, elements of one moment plane
    uint32_t rowstride;  // populations: elements of one (y,z) row block, PX * (54 or 27)
    uint32_t planeel;    // populations: elements of one z-plane, rowstride * IY (its bytes fit 32 bits)
    int y_mode, wrap_gradient, visc_mode;
    int sponge_ystart, sponge_yend;
    float nu, sponge_K;
    float taur, taub;
    float omega_b;    // 1/taub, the inlet's omegab (constants.cuh:67)
    float A_fixed;    // > 0: A used as given
    float sig4;       // sigma/(4 cs^4), constants.cuh:76
    float beta;
    float jet_x0, jet_z0, jet_R, jet_U;
    float rhor0, rhob0;
    int NR;           // radial bins, NX/2 (constants.cuh:20)
    int storage;      // MCG_STORE_*
    int inlet_profile;   // MCG_INLET_*
    float inlet_w;       // tanh width of the MCG_INLET_TANH profile
    float pert_amp;      // transverse perturbation amplitude, fraction of jet_U
    uint32_t pert_seed;  // seed of the perturbation
    int wall;            // 1: bounce-back walls on the x and z ghost layers
    float grav, ghalf;   // body force F_y = -grav rho_r on red; ghalf = grav/2
};

struct DevBufs {
    const void *f_src;     // populations read this step (Store<storage> layout)
    void *f_dst;           // populations written this step
    const float *in_src;
    float *in_dst;
    const float *out_src;
    float *out_dst;
    const float *gB_src;   // outlet ghost state read this step
    float *gB_dst;         // outlet ghost state written this step
    const float *ghost_rr;
    const float *ghost_rb;
    float *mom;            // 11 planes
    float *mom_out;        // if set: the moments of every node of this step's source state (11 planes)
    const float *inlet_du;  // if set: offset of the inlet velocity per column (3 x NX*NZ, component-major)
    unsigned long long *stats_fx;  // 7 planes of NY*NR (uy, uy2, ur, ur2, uruy, phi, phi2): this step's sums, fixed point
    unsigned long long *count;  // NY*NR
    int do_stats;
    int stats_ring;  // statistics steps: the instance with the LDS moment ring (the histogram fits beside it)
    unsigned long long step;    // index of the step being computed (0 for the first)
};

// Radial-statistics samples are summed as 64-bit two's-complement integers in
// units of 2^-36.  Integer adds give the same sum in any order, so the sums of a
// step do not depend on the order in which the atomics land; k_stats_fold then
// adds each step's sums to the FP64 running sums in a fixed order.  A summand is
// clamped to +-128 (lattice velocities and psi stay below 1, NaN goes to the
// lower bound), so a bin can take 2^20 samples per step without overflow.  The
// rounding to 2^-37 per sample is below the FP64 error of summing the samples
// one by one into sums of order 1.
constexpr double kStatScale = 68719476736.0;  // 2^36
__device__ __forceinline__ unsigned long long stat_fx(double v) {
    return (unsigned long long)__double2ll_rn(fmin(fmax(v, -128.0), 128.0) * kStatScale);
}
// one sample of node moments into a histogram slot: the 7 sums at h[q * stride],
// the count at h[7 * stride] (kernels.cu:216-232)
__device__ __forceinline__ void stat_sample(unsigned long long *h, size_t stride, double vy, double vr, double ph) {
    atomicAdd(h + 0 * stride, stat_fx(vy));
    atomicAdd(h + 1 * stride, stat_fx(vy * vy));
    atomicAdd(h + 2 * stride, stat_fx(vr));
    atomicAdd(h + 3 * stride, stat_fx(vr * vr));
    atomicAdd(h + 4 * stride, stat_fx(vr * vy));
    atomicAdd(h + 5 * stride, stat_fx(ph));
    atomicAdd(h + 6 * stride, stat_fx(ph * ph));
    atomicAdd(h + 7 * stride, 1ull);
}

// Population arrays: the 54 (FP32) or 27 (16-bit) direction rows of one (y,z)
// row lie next to each other, so one wavefront's gather touches a few
// contiguous 1 KB rows instead of 54 planes; the row blocks of a z-plane are
// contiguous, and so are the planes.  An array may exceed 2^32 elements, one
// z-plane may not: addresses are a plane base (64-bit, uniform where z is)
// plus a 32-bit offset within the plane.
//   pinplane: offset of interior column (x,y) within its plane (row 0)
//   plane_ptr: base of z-plane z of a population array
//   pidx: the element index modulo 2^32, the key of the stochastic rounding
__device__ __forceinline__ uint32_t pinplane(const DevParams &P, int x, int y) {
    return uint32_t(x - 1) + P.rowstride * uint32_t(y - 1);
}
__device__ __forceinline__ const void *plane_ptr(const DevParams &P, const void *f, int z) {
    return static_cast<const char *>(f) + size_t(P.planeel) * size_t(z - 1) * 4;
}
__device__ __forceinline__ void *plane_ptr(const DevParams &P, void *f, int z) {
    return static_cast<char *>(f) + size_t(P.planeel) * size_t(z - 1) * 4;
}
__device__ __forceinline__ uint32_t pidx(const DevParams &P, int x, int y, int z) {
    return pinplane(P, x, y) + P.planeel * uint32_t(z - 1);
}
// the same element index in 64 bits (arrays in the node layout)
__device__ __forceinline__ size_t pidx64(const DevParams &P, int x, int y, int z) {
    return size_t(P.planeel) * size_t(z - 1) + pinplane(P, x, y);
}
// moment-plane index of interior node (x,y,z)
__device__ __forceinline__ uint32_t midx(const DevParams &P, int x, int y, int z) {
    return uint32_t(x - 1) + uint32_t(P.PX) * (uint32_t(y - 1) + uint32_t(P.IY) * uint32_t(z - 1));
}
__device__ __forceinline__ size_t gidx(const DevParams &P, int x, int y, int z) {
    return size_t(x) + size_t(P.NX) * (size_t(y) + size_t(P.NY) * size_t(z));
}
// periodic wrap over the ghost layers, utilities/bounds.cuh:23-60
__device__ __forceinline__ int wrapN(int v, int n) { return v == 0 ? n - 2 : (v == n - 1 ? 1 : v); }

__device__ __forceinline__ bool is_jet(const DevParams &P, int x, int z) {  // utilities/mathUtilities.cuh:14-21
    const float dx = float(x) - P.jet_x0;
    const float dz = float(z) - P.jet_z0;
    return dx * dx + dz * dz <= P.jet_R * P.jet_R;
}

// relaxation rate of the sponge layer, utilities/mathUtilities.cuh:49-90
__device__ __forceinline__ float omega_sponge(const DevParams &P, int y) {
    const float inv_width = 1.0f / float(P.sponge_yend - P.sponge_ystart);
    const float s = clamp01((float(y) - float(P.sponge_ystart)) * inv_width);
    const float nus = P.nu * (1.0f + P.sponge_K * (s * s * s));
    const float tau = 0.5f + nus * inv_cs2;
    return 1.0f / tau;
}
// interface relaxation time, forces.cuh:77-85
__device__ __forceinline__ float tau_interface(const DevParams &P, float rr, float rb) {
    const float p = psi(rr, rb);
    return ((1.0f + p) / 2.0f) * P.taur + ((1.0f - p) / 2.0f) * P.taub;
}
__device__ __forceinline__ float omega_node(const DevParams &P, int y, float rr, float rb) {
    if (P.visc_mode == MCG_VISC_INTERFACE) return 1.0f / tau_interface(P, rr, rb);
    return omega_sponge(P, y);
}
// A = sigma/(4 cs^4 tau), forces.cuh:87-90,102-103
__device__ __forceinline__ float A_node(const DevParams &P, float rr, float rb) {
    if (P.A_fixed > 0.0f) return P.A_fixed;
    const float tau = P.visc_mode == MCG_VISC_INTERFACE ? tau_interface(P, rr, rb) : P.taub;
    return P.sig4 / tau;
}

// ---- population storage ------------------------------------------------
// Direction I of colour c sits at element (c*27 + I)*stride + off of the
// node's z-plane (stride = PX, off = pinplane; f = plane_ptr).
// FP32: two rows per direction ([colour*27 + i]).
// FP16 / BF16: one 32-bit word per direction and node: the low half holds
// t/w_i - 1 with t = f_r + f_b (the offset of the two-colour population from
// its lattice weight, scaled by 1/w_i), the high half holds the minority
// colour m/w_i -- the colour of smaller magnitude, with its own sign -- whose
// mantissa LSB is replaced by the colour flag (1: red is the minority, 0:
// blue).  The decode clears the flag bit before converting, so a minority of
// 0 stays exactly 0: pure phases are exact also where their populations are
// negative, and a negative minority keeps its sign.  The minority keeps 9
// (FP16) or 6 (BF16) mantissa bits; the arithmetic stays FP32.
constexpr float invw(int i) { return 1.0f / w(i); }

__device__ __forceinline__ uint32_t f32_bits(float v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ float bits_f32(uint32_t v) { return __builtin_bit_cast(float, v); }

template <int ST>
struct Store;

template <>
struct Store<MCG_STORE_FP32> {
    static constexpr int planes = 2 * Q;
    // raw words of one node's gather: load_raw issues the loads, decode_raw
    // turns the words of direction I into (f_r, f_b) later
    static constexpr int kWords = 2;
    template <int I>
    __device__ __forceinline__ static void load_raw(const void *f, uint32_t stride, uint32_t off, uint32_t *w) {
        const uint32_t *p = static_cast<const uint32_t *>(f);
        w[0] = p[uint32_t(I) * stride + off];
        w[1] = p[uint32_t(Q + I) * stride + off];
    }
    template <int I>
    __device__ __forceinline__ static void decode_raw(const uint32_t *w, float &fr, float &fb, float &g) {
        fr = __builtin_bit_cast(float, w[0]);
        fb = __builtin_bit_cast(float, w[1]);
        g = fr + fb;
    }
    template <int I>
    __device__ __forceinline__ static void load(const void *f, uint32_t stride, uint32_t off, float &fr, float &fb,
                                                float &g) {
        const float *p = static_cast<const float *>(f);
        fr = p[uint32_t(I) * stride + off];
        fb = p[uint32_t(Q + I) * stride + off];
        g = fr + fb;
    }
    template <int I>
    __device__ __forceinline__ static void load(const void *f, uint32_t stride, uint32_t off, float &fr, float &fb) {
        float g;
        load<I>(f, stride, off, fr, fb, g);
    }
    // population stores are streaming (non-temporal): nothing reads them back
    // within the step, and keeping them out of L2 leaves it to the gathered rows
    // that neighbouring tiles read again (-15% fetched bytes, +12% MLUPS)
    template <int I>
    __device__ __forceinline__ static void store(void *f, uint32_t stride, uint32_t off, float fr, float fb) {
        float *p = static_cast<float *>(f);
        __builtin_nontemporal_store(fr, &p[uint32_t(I) * stride + off]);
        __builtin_nontemporal_store(fb, &p[uint32_t(Q + I) * stride + off]);
    }
};

template <int ST>
struct Store16 {
    static constexpr int planes = Q;
    // word idx of a z-plane as a 32-bit byte offset from the plane's base
    // (plane_ptr; a plane stays below 4 GiB), so the loads and stores address
    // base (scalar) + offset (32-bit vector register)
    __device__ __forceinline__ static const uint32_t *at(const void *f, uint32_t idx) {
        return reinterpret_cast<const uint32_t *>(static_cast<const char *>(f) + (idx << 2));
    }
    __device__ __forceinline__ static uint32_t *at(void *f, uint32_t idx) {
        return reinterpret_cast<uint32_t *>(static_cast<char *>(f) + (idx << 2));
    }
    __device__ __forceinline__ static float half_value(uint32_t h) {
        if constexpr (ST == MCG_STORE_FP16) {
            return float(__builtin_bit_cast(_Float16, uint16_t(h)));
        } else {
            return bits_f32(h << 16);
        }
    }
    static constexpr uint32_t kFlag = 0x10000u;  // LSB of the high half: red is the minority
    // t = (a + 1) w as one fused multiply-add; the colours m and t - m; g = t
    template <int I>
    __device__ __forceinline__ static void decode(uint32_t v, float &fr, float &fb, float &g) {
        const float a = half_value(v & 0xffffu);
        const float t = __builtin_fmaf(a, w(I), w(I));
        const float m = half_value((v >> 16) & 0xfffeu) * w(I);
        if (v & kFlag) {
            fr = m;
            fb = t - m;
        } else {
            fb = m;
            fr = t - m;
        }
        g = t;
    }
    // the flag bit inserted into a word whose high half is already rounded
    // to an even mantissa (v_bfi_b32)
    __device__ __forceinline__ static uint32_t with_flag(uint32_t v, bool red_minor) {
        return (v & ~kFlag) | (red_minor ? kFlag : 0u);
    }
    // The word from a = t/w - 1 and the signed minority m/w, both rounded to
    // nearest: the low half by the convert (ties to even), the high half on
    // the grid of its even mantissas (ties away from zero): half of that grid
    // step (2^13 of the FP32 pattern for FP16, 2^16 for BF16) is added before
    // the truncation drops 14 (FP16) or 17 (BF16) bits.
    __device__ __forceinline__ static uint32_t pack_rn(float lo, float hi, bool red_minor) {
        uint32_t l, h;
        if constexpr (ST == MCG_STORE_FP16) {
            l = __builtin_bit_cast(uint16_t, _Float16(lo));
            const uint32_t hb = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(bits_f32(f32_bits(hi) + 0x2000u), 0.0f));
            h = hb & 0xfffeu;
        } else {
            l = __builtin_bit_cast(uint16_t, __bf16(lo));
            h = ((f32_bits(hi) + 0x10000u) >> 16) & 0xfffeu;
        }
        return l | (h << 16) | (red_minor ? kFlag : 0u);
    }
    template <int I>
    __device__ __forceinline__ static uint32_t encode(float fr, float fb) {
        const float a = (fr + fb) * invw(I) - 1.0f;
        const bool red_minor = fabsf(fr) < fabsf(fb);
        return pack_rn(a, (red_minor ? fr : fb) * invw(I), red_minor);
    }
    template <int I>
    __device__ __forceinline__ static void load(const void *f, uint32_t stride, uint32_t off, float &fr, float &fb,
                                                float &g) {
        decode<I>(*at(f, uint32_t(I) * stride + off), fr, fb, g);
    }
    template <int I>
    __device__ __forceinline__ static void load(const void *f, uint32_t stride, uint32_t off, float &fr, float &fb) {
        float g;
        load<I>(f, stride, off, fr, fb, g);
    }
    static constexpr int kWords = 1;
    template <int I>
    __device__ __forceinline__ static void load_raw(const void *f, uint32_t stride, uint32_t off, uint32_t *w) {
        w[0] = *at(f, uint32_t(I) * stride + off);
    }
    template <int I>
    __device__ __forceinline__ static void decode_raw(const uint32_t *w, float &fr, float &fb, float &g) {
        decode<I>(w[0], fr, fb, g);
    }
    template <int I>
    __device__ __forceinline__ static void store(void *f, uint32_t stride, uint32_t off, float fr, float fb) {
        __builtin_nontemporal_store(encode<I>(fr, fb), at(f, uint32_t(I) * stride + off));
    }
    // The same word with stochastic rounding: a uniform random integer below
    // the dropped bits is added to the FP32 bit pattern, and the result is
    // truncated toward zero.  The pattern is sign-magnitude and ordered by
    // magnitude, so the magnitude rounds up with probability equal to the
    // dropped fraction and is truncated otherwise: E[stored] = value, and the
    // sign is kept (also for -0).  The low half drops 13 (FP16) or 16 (BF16)
    // bits, the high half one more, the bit the flag then takes.  A steady
    // flow thus carries zero-mean storage noise instead of the same rounding
    // error every step.  FP16: one packed round-toward-zero convert
    // (v_cvt_pkrtz_f16_f32) of both halves; below the FP16 normal range
    // (|x| < 2^-14) more bits drop and the rounding leans toward zero by less
    // than 2^-24.  BF16: the two high halves are the word (one byte permute).
    // Integer adds are used because the v_cvt_sr_* converts ran the FP16 step
    // 5% slower.
    __device__ __forceinline__ static uint32_t pack_sr(float lo, float hi, uint32_t r, bool red_minor) {
        uint32_t v;
        if constexpr (ST == MCG_STORE_FP16) {
            v = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(bits_f32(f32_bits(lo) + (r >> 19)),
                                                                        bits_f32(f32_bits(hi) + (r >> 18))));
        } else {
            // r >> 16 as a byte permute (selector 0x0c: a zero byte): the
            // compiler would otherwise fold the shift into word-select (SDWA)
            // adds, which ran the BF16 step 4.5% slower
            const uint32_t d = __builtin_amdgcn_perm(0u, r, 0x0c0c0302u);
            // bytes 2,3 of (lo + d), then bytes 2,3 of (hi + (r >> 15))
            v = __builtin_amdgcn_perm(f32_bits(hi) + (r >> 15), f32_bits(lo) + d, 0x07060302u);
        }
        return with_flag(v, red_minor);
    }
    // word from the offset a = t/w - 1 and the colours already divided by w,
    // stochastically rounded with the random word r (both halves)
    template <int I>
    __device__ __forceinline__ static void store_w(void *f, uint32_t stride, uint32_t off, float a, float rw, float bw,
                                                   uint32_t r) {
        const bool red_minor = fabsf(rw) < fabsf(bw);
        __builtin_nontemporal_store(pack_sr(a, red_minor ? rw : bw, r, red_minor), at(f, uint32_t(I) * stride + off));
    }
};
template <>
struct Store<MCG_STORE_FP16> : Store16<MCG_STORE_FP16> {};
template <>
struct Store<MCG_STORE_BF16> : Store16<MCG_STORE_BF16> {};

// Source of the population that streams into interior node (x,y,z) along
// direction I: kind 0 = interior node (z = its plane, off = its offset within
// the plane), 1 / 2 = the inlet / outlet source buffer (off = element of colour
// r, colour b at +NX*NZ), 3 = bounce-back at an x/z wall (the node itself,
// direction opp(I)).
struct Src {
    int kind;
    int z;
    size_t off;
};

template <int I>
__device__ __forceinline__ Src source_of(const DevParams &P, int x, int y, int z) {
    int xs = x - kCX[I];
    int ys = y - kCY[I];
    int zs = z - kCZ[I];
    // half-way bounce-back: a population arriving from a wall ghost is the
    // node's own post-collision population of the opposite direction
    if (P.wall && (xs < 1 || xs > P.NX - 2 || zs < 1 || zs > P.NZ - 2)) return Src{3, z, size_t(pinplane(P, x, y))};
    if constexpr (kCY[I] != 0) {
        if (P.y_mode == MCG_Y_JET) {
            const bool ghost = kCY[I] > 0 ? (ys == 0) : (ys == P.NY - 1);
            if (ghost) {
                // ghost-plane sources are not wrapped (jet/boundary_conditions.cuh:73,222)
                const size_t nxz = size_t(P.NX) * size_t(P.NZ);
                const size_t off = size_t(slot_of(I)) * 2 * nxz + size_t(zs) * P.NX + size_t(xs);
                return Src{kCY[I] > 0 ? 1 : 2, 0, off};
            }
        } else {
            ys = wrapN(ys, P.NY);
        }
    }
    xs = wrapN(xs, P.NX);
    zs = wrapN(zs, P.NZ);
    return Src{0, zs, size_t(pinplane(P, xs, ys))};
}

// g: f_r + f_b as the moments sum it (Store::load)
template <int ST, int I>
__device__ __forceinline__ void gather(const DevParams &P, const void *__restrict__ f, const float *__restrict__ bin,
                                       const float *__restrict__ bout, int x, int y, int z, float &fr, float &fb,
                                       float &g) {
    const Src s = source_of<I>(P, x, y, z);
    if (s.kind == 0) {
        Store<ST>::template load<I>(plane_ptr(P, f, s.z), uint32_t(P.PX), uint32_t(s.off), fr, fb, g);
    } else if (s.kind == 3) {
        Store<ST>::template load<opp(I)>(plane_ptr(P, f, s.z), uint32_t(P.PX), uint32_t(s.off), fr, fb, g);
    } else {
        const float *b = s.kind == 1 ? bin : bout;
        fr = b[s.off];
        fb = b[s.off + size_t(P.NX) * P.NZ];
        g = fr + fb;
    }
}
template <int ST, int I>
__device__ __forceinline__ void gather(const DevParams &P, const void *__restrict__ f, const float *__restrict__ bin,
                                       const float *__restrict__ bout, int x, int y, int z, float &fr, float &fb) {
    float g;
    gather<ST, I>(P, f, bin, bout, x, y, z, fr, fb, g);
}

template <int ST, int I>
__device__ __forceinline__ void scatter(const DevParams &P, void *f, float *bin, float *bout, int x, int y, int z,
                                        float fr, float fb) {
    const Src s = source_of<I>(P, x, y, z);
    if (s.kind == 0) {
        Store<ST>::template store<I>(plane_ptr(P, f, s.z), uint32_t(P.PX), uint32_t(s.off), fr, fb);
    } else if (s.kind == 3) {
        Store<ST>::template store<opp(I)>(plane_ptr(P, f, s.z), uint32_t(P.PX), uint32_t(s.off), fr, fb);
    } else {
        float *b = s.kind == 1 ? bin : bout;
        b[s.off] = fr;
        b[s.off + size_t(P.NX) * P.NZ] = fb;
    }
}

// moments of interior node (x,y,z) from the gathered populations (lbm.cuh:82-165)
template <int ST>
__device__ __forceinline__ Moments gather_moments(const DevParams &P, const void *__restrict__ f,
                                                  const float *__restrict__ bin, const float *__restrict__ bout,
                                                  int x, int y, int z) {
    MomentAcc acc;
    unroll<0, Q>([&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        float a, b, g;
        gather<ST, I>(P, f, bin, bout, x, y, z, a, b, g);
        acc.add<I>(a, b, g);
    });
    return acc.finish(P.ghalf);
}

__device__ __forceinline__ bool interior(const DevParams &P, int x, int y, int z) {
    return x >= 1 && x <= P.NX - 2 && y >= 1 && y <= P.NY - 2 && z >= 1 && z <= P.NZ - 2;
}

// psi at a ghost node, as the colour gradient reads it (forces.cuh:31-33):
// the outlet plane's inner nodes hold the convective state of the previous
// step, every other ghost node its ghost density.
__device__ __forceinline__ float ghost_psi(const DevParams &P, const DevBufs &B, int x, int y, int z) {
    if (P.y_mode == MCG_Y_JET && y == P.NY - 1 && x >= 1 && x <= P.NX - 2 && z >= 1 && z <= P.NZ - 2) {
        const size_t nxz = size_t(P.NX) * P.NZ;
        const size_t o = size_t(z) * P.NX + x;
        return psi(B.gB_src[o], B.gB_src[nxz + o]);
    }
    const size_t g = gidx(P, x, y, z);
    return psi(B.ghost_rr[g], B.ghost_rb[g]);
}

// neighbour coordinates for the colour gradient (wrapped when wrap_gradient)
__device__ __forceinline__ void grad_neighbor(const DevParams &P, int x, int y, int z, int &xn, int &yn, int &zn) {
    xn = x; yn = y; zn = z;
    if (P.wrap_gradient) {  // along the periodic axes
        if (!P.wall) {
            xn = wrapN(x, P.NX);
            zn = wrapN(z, P.NZ);
        }
        if (P.y_mode == MCG_Y_PERIODIC) yn = wrapN(y, P.NY);
    }
}

// Per-node collision set-up, lbm.cuh:193-250 and forces.cuh:92-105; the
// colour gradient is evaluated at interface nodes only, as the reference does.
template <class PhiFn>
__device__ __forceinline__ NodeColl node_setup(const DevParams &P, const Moments &m, int y, PhiFn &&phi) {
    NodeColl n;
    const float rT = m.rr + m.rb;
    const float invrT = 1.0f / rT;
    n.aR = m.rr * invrT;
    n.aB = m.rb * invrT;
    const float oms = 1.0f - omega_node(P, y, m.rr, m.rb);
    n.ux = m.ux;
    n.uy = m.uy;
    n.uz = m.uz;
    const float usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
    n.e0 = rT * (1.0f - usq * inv_2cs2);
    n.force = P.grav != 0.0f;
    n.g9 = 0.0f;
    n.g3 = 0.0f;
    if (n.force) {
        // body force on red, F_y = -grav rho_r: even source (1 - omega/2)(9 (c.u) c_y - 3 u_y) F_y,
        // odd 1.5 c_y F_y (the regularised collision drops the first-order non-equilibrium
        // moment -F/2, whose (1 - omega) share the odd source restores: sum c f* = j + F)
        const float Fy = -P.grav * m.rr;
        const float gy = (0.5f + 0.5f * oms) * Fy;
        n.g9 = 9.0f * gy;
        n.g3 = 1.5f * Fy;
        n.e0 = fma_(-3.0f * gy, m.uy, n.e0);
    }
    n.e1 = rT * inv_2cs4;
    n.o1 = rT * (inv_cs2 - usq * inv_2cs4);
    n.o3 = rT * inv_6cs6;
    n.q2 = oms * inv_2cs4;
    n.q3 = oms * inv_2cs6;
    n.pxx = m.pxx;
    n.pyy = m.pyy;
    n.pzz = m.pzz;
    n.dxy = 2.0f * m.pxy;
    n.dyz = 2.0f * m.pyz;
    n.dxz = 2.0f * m.pxz;
    n.tr = -cs2 * (m.pxx + m.pyy + m.pzz);
    n.txxy = m.pxx * m.uy + n.dxy * m.ux;
    n.txxz = m.pxx * m.uz + n.dxz * m.ux;
    n.txyy = m.pyy * m.ux + n.dxy * m.uy;
    n.txzz = m.pzz * m.ux + n.dxz * m.uz;
    n.tyyz = m.pyy * m.uz + n.dyz * m.uy;
    n.tyzz = m.pzz * m.uy + n.dyz * m.uz;
    n.txyz = n.dxy * m.uz + n.dxz * m.uy + n.dyz * m.ux;
    n.Fx = n.Fy = n.Fz = n.s2 = n.sB = n.rc = 0.0f;
    const float I = 4.0f * m.rr * m.rb * invrT * invrT;
    if (!(I <= 1.0e-4f)) {
        color_gradient(phi, n.Fx, n.Fy, n.Fz);
        const float hA = 0.5f * A_node(P, m.rr, m.rb);
        const float absF = sqrtf(n.Fx * n.Fx + n.Fy * n.Fy + n.Fz * n.Fz);
        n.sB = hA * absF;
        if (!(absF <= 1.0e-4f)) {
            const float invF = 1.0f / absF;
            n.s2 = hA * invF;
            n.rc = P.beta * ((m.rr * m.rb) / (rT * rT)) * rT * invF;
        }
    }
    return n;
}

// All 54 post-collision populations of a node, handed to put(I, f_r, f_b).
template <bool BULK = false, class PutFn>
__device__ __forceinline__ void collide_node(const NodeColl &n, PutFn &&put) {
    float r0, b0;
    collide_rest<BULK>(n, r0, b0);
    put(std::integral_constant<int, 0>{}, r0, b0);
    unroll<0, 13>([&](auto kc) {
        constexpr int I = 2 * decltype(kc)::value + 1;
        float rp, bp, rm, bm;
        collide_pair<I, BULK>(n, rp, bp, rm, bm);
        put(std::integral_constant<int, I>{}, rp, bp);
        put(std::integral_constant<int, I + 1>{}, rm, bm);
    });
}

// lowbias32, an integer hash with full avalanche (C. Wellons,
// "Prospecting for hash functions", 2018: xorshift-multiply rounds 16,
// 0x7feb352d, 15, 0x846ca68b, 16)
__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// Random word of the stochastic rounding of node p's populations at a step:
// a function of (p, step) only, so every kernel that stores node p at that
// step draws the same bits and the results stay reproducible.
__device__ __forceinline__ uint32_t sr_step_key(uint64_t step) { return lowbias32(uint32_t(step) ^ 0x2545F491u); }
__device__ __forceinline__ uint32_t sr_node_key(uint32_t p, uint32_t step_key) { return lowbias32(p ^ step_key); }

// Collide node n and store its 54 post-collision populations at element p of
// the row block (FP32 rows, or 16-bit words built directly from the w-scaled
// terms, stochastically rounded).  Random words: h = sr_node_key(p, step_key)
// for the rest population, h + k G for the opposite pair (2k-1, 2k), both
// halves of both words, G = 0x9E3779B9 (a Weyl sequence: distinct high bits
// for the 14 groups).  Every stored half is unbiased whatever the sharing;
// sharing only correlates the noise within a node.
// f: the node's z-plane (plane_ptr), p: its offset there, key: its element
// index (pidx), which keys the stochastic rounding
template <int ST, bool BULK>
__device__ __forceinline__ void collide_store_as(const NodeColl &n, void *f, uint32_t PX, uint32_t p, uint32_t key,
                                                 uint32_t step_key) {
    if constexpr (ST == MCG_STORE_FP32) {
        collide_node<BULK>(n, [&](auto Ic, float a, float b) {
            Store<ST>::template store<decltype(Ic)::value>(f, PX, p, a, b);
        });
    } else {
        const uint32_t h = sr_node_key(key, step_key);
        {
            float E = fma_(n.q2, n.tr, n.e0);
            if constexpr (!BULK) E = fma_(n.sB, 1.25f, E);
            Store<ST>::template store_w<0>(f, PX, p, E - 1.0f, n.aR * E, n.aB * E, h);
        }
        unroll<0, 13>([&](auto kc) {
            constexpr int I = 2 * decltype(kc)::value + 1;
            float ap, rp, bp, am, rm, bm;
            collide_pair_w<I, BULK>(n, ap, rp, bp, am, rm, bm);
            const uint32_t r = h + uint32_t(decltype(kc)::value + 1) * 0x9E3779B9u;
            Store<ST>::template store_w<I>(f, PX, p, ap, rp, bp, r);
            Store<ST>::template store_w<I + 1>(f, PX, p, am, rm, bm, r);
        });
    }
}

// FP32: a wave whose nodes all carry no Omega2 or recolouring (bulk nodes, as
// most are away from the interface) takes the form without those terms, about
// 80 fewer operations per node and the same values (256x512x256 jet 1.3%
// faster, 192x384x192 rising bubble 4.5%).  The 16-bit steps keep the one
// general form: the second copy of the collision ran them 1.5-4% slower.
template <int ST>
__device__ __forceinline__ void collide_store(const NodeColl &n, void *f, uint32_t PX, uint32_t p, uint32_t key,
                                              uint32_t step_key) {
    if constexpr (ST == MCG_STORE_FP32) {
        const bool bulk = n.s2 == 0.0f && n.sB == 0.0f && n.rc == 0.0f;
        if (__all(bulk)) {
            collide_store_as<ST, true>(n, f, PX, p, key, step_key);
            return;
        }
    }
    collide_store_as<ST, false>(n, f, PX, p, key, step_key);
}

// SplitMix64 (Steele, Lea and Flood 2014, the finaliser constants of Vigna's
// reference implementation): a counter-based generator, so the CPU and the GPU
// draw the same numbers for the same (seed, step, node).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform number in [0,1) (24 bits) number c of inlet column (x,z) at step n
__device__ __forceinline__ float inlet_xi(uint32_t seed, uint64_t step, int x, int z, int NX, int c) {
    const uint64_t key = splitmix64(splitmix64(splitmix64(uint64_t(seed)) ^ step) ^ (uint64_t(z) * uint64_t(NX) + uint64_t(x)));
    return float(splitmix64(key + uint64_t(c)) >> 40) * (1.0f / 16777216.0f);
}

// Inlet velocity of jet ghost node (x,0,z).  Sharp profile: (0, U, 0)
// (jet/boundary_conditions.cuh:45-47).  MCG_INLET_TANH (an extension beyond
// the reference): u_y = U (1 - tanh((r - R)/w))/2 on the jet disk, and
// transverse u_x, u_z = amp U (2 xi - 1) with xi drawn per step and column.
__device__ __forceinline__ void inlet_velocity(const DevParams &P, const float *du, uint64_t step, int x, int z,
                                               float &ux, float &uy, float &uz) {
    ux = 0.0f;
    uy = P.jet_U;
    uz = 0.0f;
    if (P.inlet_profile == MCG_INLET_TANH) {
        const float dx = float(x) - P.jet_x0;
        const float dz = float(z) - P.jet_z0;
        const float r = sqrtf(dx * dx + dz * dz);
        uy = P.jet_U * (0.5f * (1.0f - tanhf((r - P.jet_R) / P.inlet_w)));
        const float a = P.pert_amp * P.jet_U;
        ux = a * (2.0f * inlet_xi(P.pert_seed, step, x, z, P.NX, 1) - 1.0f);
        uz = a * (2.0f * inlet_xi(P.pert_seed, step, x, z, P.NX, 2) - 1.0f);
    }
    if (du) {  // control of the inverse jet problem (an extension)
        const size_t nxz = size_t(P.NX) * P.NZ, o = size_t(z) * P.NX + x;
        ux += du[o];
        uy += du[nxz + o];
        uz += du[2 * nxz + o];
    }
}

// Inlet populations sent by ghost node (x,0,z) of a jet column, computed from
// the moments m of node (x,1,z): jet/boundary_conditions.cuh:17-85.
__device__ __forceinline__ void inlet_emit(const DevParams &P, float *__restrict__ in_dst, const Moments &m, int x, int z,
                                           uint64_t step, const float *du) {
    const float rrB = (1.0f - 1.0f) * P.rhor0;  // jet node: is_jet = 1
    const float rbB = 1.0f * P.rhob0;
    const float rT = rrB + rbB;
    const float aR = rrB * (1.0f / rT);
    const float aB = rbB * (1.0f / rT);
    float om = P.omega_b;
    if (P.visc_mode == MCG_VISC_INTERFACE) om = 1.0f / tau_interface(P, rrB, rbB);
    const float oms = 1.0f - om;
    float ex, ey, ez;
    inlet_velocity(P, du, step, x, z, ex, ey, ez);
    const size_t nxz = size_t(P.NX) * P.NZ;
    const size_t o = size_t(z) * P.NX + x;
    unroll<0, 9>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int I = kUp[k];
        const float g = boundary_pop<I>(rT, ex, ey, ez, oms, m);
        in_dst[(2 * k) * nxz + o] = aR * g;
        in_dst[(2 * k + 1) * nxz + o] = aB * g;
    });
}

// Convective outlet for column (x,z) from the moments m of node (x,NY-2,z):
// jet/boundary_conditions.cuh:163-233.
__device__ __forceinline__ void outlet_emit(const DevParams &P, const float *__restrict__ gB_src, float *__restrict__ gB_dst,
                                            float *__restrict__ out_dst, const Moments &m, int x, int z) {
    const size_t nxz = size_t(P.NX) * P.NZ;
    const size_t o = size_t(z) * P.NX + x;
    const float uc = clamp01(m.uy);
    const float rrB = gB_src[o] - uc * (gB_src[o] - m.rr);
    const float rbB = gB_src[nxz + o] - uc * (gB_src[nxz + o] - m.rb);
    const float uxB = gB_src[2 * nxz + o] - uc * (gB_src[2 * nxz + o] - m.ux);
    const float uyB = gB_src[3 * nxz + o] - uc * (gB_src[3 * nxz + o] - m.uy);
    const float uzB = gB_src[4 * nxz + o] - uc * (gB_src[4 * nxz + o] - m.uz);
    gB_dst[o] = rrB;
    gB_dst[nxz + o] = rbB;
    gB_dst[2 * nxz + o] = uxB;
    gB_dst[3 * nxz + o] = uyB;
    gB_dst[4 * nxz + o] = uzB;
    const float rhoT = rrB + rbB;
    const float inv = 1.0f / rhoT;
    const float aR = rrB * inv;
    const float aB = rbB * inv;
    const float oms = 1.0f - omega_node(P, P.NY - 2, rrB, rbB);
    unroll<0, 9>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int I = kDown[k];
        const float g = boundary_pop<I>(rhoT, uxB, uyB, uzB, oms, m);
        out_dst[(2 * k) * nxz + o] = aR * g;
        out_dst[(2 * k + 1) * nxz + o] = aB * g;
    });
}

// Radial bin of interior column (x,z), kernels.cu:207-215; -1 when outside.
__device__ __forceinline__ int radial_bin(const DevParams &P, int x, int z, float &dx, float &dz, float &r) {
    dx = float(x) - P.jet_x0;
    dz = float(z) - P.jet_z0;
    r = sqrtf(dx * dx + dz * dz);
    const uint32_t rb = uint32_t(r);
    return rb >= uint32_t(P.NR) ? -1 : int(rb);
}

}  // namespace mcg
