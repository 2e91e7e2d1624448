// schedule.hpp -- banded launch schedules of the step kernel (host code, no
// HIP types: the unit tables are built on the host and read by the kernel).
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

namespace mcg {

// One workgroup's work in a banded schedule: a tile (x-tile + nxt * y-block)
// and its planes zb..ze (1-based); ze < zb marks padding (the workgroup
// leaves at once).
struct FusedUnit {
    uint32_t tile, zb, ze, pad;
};

// ---- banded schedules ---------------------------------------------------
// Workgroup b runs on XCD b mod 8 and, within its XCD, starts on the first CU
// that frees up (one workgroup per CU).  With units of one length the last
// round is partial: the 16-bit step at 256x512x256 ran 1,204 units of 37
// planes on 256 CUs (4.7 rounds), and 30% of the CUs sat idle through the
// last round, 10% of the step.  A banded schedule deals the tile-planes (tile
// major) to the XCDs in equal contiguous shares, so a share is a run of
// neighbouring tiles, cuts each share at z-band edges of decreasing thickness,
// and lists an XCD's units band by band: the long units run first and the
// short ones fill the CUs that finish early.  Each unit still costs its fill
// planes, so the thickness rule is chosen by an estimate of the span
// (schedule_span) and the launch-shape timing decides among the best three
// and two other dealings of the best (BandDeal).
constexpr int kXcds = 8;
constexpr int kBandedCandidates = 3;

// band thicknesses A r^j, at least tmin each, the last band taking the rest
inline std::vector<int> band_thickness(int IZ, double A, double r, int tmin) {
    std::vector<int> th;
    int z = 0;
    double a = A;
    while (z < IZ) {
        int t = std::max(tmin, int(a + 0.5));
        if (IZ - z - t < tmin) t = IZ - z;
        th.push_back(t);
        z += t;
        a *= r;
    }
    return th;
}

// How the tile-planes (tile major) are dealt to the XCDs: kDealContiguous,
// one equal contiguous share each; kDealPaired, 16 equal chunks, XCD k taking
// chunks k and k+8; kDealMirrored, XCD k taking chunks k and 15-k.  The cost
// of a tile-plane is not uniform: at 256x512x256 the 16-bit XCD holding the
// jet's inlet tiles was busy 6% longer than the one holding the outlet tiles
// (half of whose rows lie past the grid), and the FP32 XCDs at both ends of y
// 4% longer than those in the middle; two chunks from different places even
// such profiles out, at the price of more tile boundaries inside an XCD.
enum BandDeal { kDealContiguous = 0, kDealPaired = 1, kDealMirrored = 2 };

// the units of a band rule, in launch order (XCD k at indexes k, k + 8, ...;
// XCDs with fewer units padded with empty ones); planes are 1-based
inline std::vector<FusedUnit> banded_units(int ntiles, int IZ, const std::vector<int> &th,
                                           BandDeal deal = kDealContiguous) {
    const int64_t W = int64_t(ntiles) * IZ;
    const int nchunk = deal == kDealContiguous ? kXcds : 2 * kXcds;
    std::vector<std::vector<FusedUnit>> per(kXcds);
    for (int k = 0; k < kXcds; ++k) {
        int chunks[2] = {k, -1};
        if (deal == kDealPaired) chunks[1] = k + kXcds;
        if (deal == kDealMirrored) chunks[1] = 2 * kXcds - 1 - k;
        int b0 = 0;
        for (int t : th) {
            const int b1 = b0 + t;
            for (int c : chunks) {
                if (c < 0) continue;
                const int64_t s0 = W * c / nchunk, s1 = W * (c + 1) / nchunk;
                if (s1 <= s0) continue;
                const int t0 = int(s0 / IZ), t1 = int((s1 - 1) / IZ);
                for (int tile = t0; tile <= t1; ++tile) {
                    const int za = int(std::max<int64_t>(0, s0 - int64_t(tile) * IZ));
                    const int zb = int(std::min<int64_t>(IZ, s1 - int64_t(tile) * IZ));
                    const int lo = std::max(za, b0), hi = std::min(zb, b1);
                    if (hi > lo) per[k].push_back(FusedUnit{uint32_t(tile), uint32_t(lo + 1), uint32_t(hi), 0u});
                }
            }
            b0 = b1;
        }
    }
    size_t n = 0;
    for (const auto &u : per) n = std::max(n, u.size());
    std::vector<FusedUnit> out;
    out.reserve(n * kXcds);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < kXcds; ++k) out.push_back(i < per[k].size() ? per[k][i] : FusedUnit{0u, 1u, 0u, 0u});
    return out;
}

// Estimated span of a launch, in planes: per XCD, cus CUs take the units in
// order, each costing its planes plus `fill` (the fill iterations and the
// start of a workgroup: about 1.9 planes with 16-bit storage, 1.5 with FP32).
inline double schedule_span(const std::vector<FusedUnit> &units, int cus, double fill) {
    double span = 0.0;
    for (int k = 0; k < kXcds; ++k) {
        // free times of the XCD's CUs, earliest on top
        std::priority_queue<double, std::vector<double>, std::greater<double>> cu;
        for (int c = 0; c < std::max(1, cus); ++c) cu.push(0.0);
        double end = 0.0;
        for (size_t i = size_t(k); i < units.size(); i += kXcds) {
            const FusedUnit &u = units[i];
            if (u.ze < u.zb) continue;
            const double t = cu.top() + double(u.ze - u.zb + 1) + fill;
            cu.pop();
            cu.push(t);
            end = std::max(end, t);
        }
        span = std::max(span, end);
    }
    return span;
}

struct BandedSchedule {
    int bands;
    std::vector<int> thickness;    // of the z-bands, in order
    BandDeal deal;
    std::vector<FusedUnit> units;  // in launch order
};

// The kBandedCandidates band rules of shortest estimated span among A = IZ/k
// (k = 1..12), r = 0.2..1.0 and tmin = 2..16, with 2 to 64 bands and at most
// 16384 units, for ntiles tiles of IZ planes on kXcds XCDs of cus CUs, dealt
// in contiguous shares; then the best of them dealt paired and mirrored.
inline std::vector<BandedSchedule> banded_rules(int ntiles, int IZ, int cus, double fill) {
    std::vector<std::pair<double, std::vector<int>>> rules;
    for (int k = 1; k <= 12; ++k)
        for (int ri = 2; ri <= 10; ++ri)
            for (int tmin : {2, 3, 4, 6, 8, 12, 16}) {
                std::vector<int> th = band_thickness(IZ, double(IZ) / k, 0.1 * ri, tmin);
                if (th.size() < 2 || th.size() > 64) continue;
                bool seen = false;
                for (const auto &r : rules) seen = seen || r.second == th;
                if (seen) continue;
                const std::vector<FusedUnit> u = banded_units(ntiles, IZ, th);
                if (u.size() > 16384) continue;
                rules.emplace_back(schedule_span(u, cus, fill), th);
            }
    std::stable_sort(rules.begin(), rules.end(),
                     [](const auto &a, const auto &b) { return a.first < b.first; });
    std::vector<BandedSchedule> out;
    for (size_t i = 0; i < rules.size() && int(out.size()) < kBandedCandidates; ++i) {
        const std::vector<int> &th = rules[i].second;
        out.push_back(BandedSchedule{int(th.size()), th, kDealContiguous, banded_units(ntiles, IZ, th)});
    }
    // the best rule also with the tile-planes dealt in two chunks per XCD
    if (!out.empty()) {
        const std::vector<int> th = out.front().thickness;
        for (BandDeal d : {kDealPaired, kDealMirrored})
            out.push_back(BandedSchedule{int(th.size()), th, d, banded_units(ntiles, IZ, th, d)});
    }
    return out;
}


}  // namespace mcg
