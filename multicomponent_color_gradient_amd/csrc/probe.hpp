// probe.hpp -- plain streaming kernels that measure the HBM bandwidth the
// device reaches for a read-only, a write-only and a copy stream (no
// counterpart in the reference; the ceiling the step kernels are compared
// with).  Each lane moves U vectors of VW dwords per pass of a grid-stride
// loop: U independent loads in flight, then their use.  Stores may be
// non-temporal (streaming past L2), as the step kernels' population stores
// are.  The read-only kernel folds what it loads into one word per lane and
// writes it only if it equals a key the host never produces, so the loads
// cannot be dropped and nothing is written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcg {

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 4;

enum ProbeKind { kProbeRead = 0, kProbeWrite = 1, kProbeCopy = 2 };

template <int KIND, int VW, bool NT>
__global__ __launch_bounds__(kProbeThreads) void k_probe(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                         size_t nvec, uint32_t key, uint32_t *__restrict__ sink) {
    typedef uint32_t V __attribute__((ext_vector_type(VW)));
    const V *s = reinterpret_cast<const V *>(src);
    V *d = reinterpret_cast<V *>(dst);
    const size_t stride = size_t(gridDim.x) * kProbeThreads * kProbeUnroll;
    uint32_t acc = 0;
    for (size_t i = size_t(blockIdx.x) * kProbeThreads * kProbeUnroll + threadIdx.x; i < nvec; i += stride) {
        V v[kProbeUnroll];
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) {
            const size_t j = i + size_t(u) * kProbeThreads;
            if constexpr (KIND == kProbeWrite) {
                v[u] = V(key + uint32_t(j));
            } else {
                v[u] = j < nvec ? s[j] : V(0u);
            }
        }
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) {
            const size_t j = i + size_t(u) * kProbeThreads;
            if constexpr (KIND == kProbeRead) {
#pragma unroll
                for (int e = 0; e < VW; ++e) acc ^= v[u][e];
            } else if (j < nvec) {
                if constexpr (NT) {
                    __builtin_nontemporal_store(v[u], &d[j]);
                } else {
                    d[j] = v[u];
                }
            }
        }
    }
    if constexpr (KIND == kProbeRead) {
        if (acc == key) sink[0] = acc;
    }
}

// launch one probe variant on the null stream (vec_bytes 4, 8 or 16)
template <int KIND, int VW>
inline void launch_probe(bool nt, int blocks, const uint32_t *src, uint32_t *dst, size_t nvec, uint32_t *sink) {
    if (nt) {
        hipLaunchKernelGGL((k_probe<KIND, VW, true>), dim3(blocks), dim3(kProbeThreads), 0, nullptr, src, dst, nvec,
                           0x9E3779B9u, sink);
    } else {
        hipLaunchKernelGGL((k_probe<KIND, VW, false>), dim3(blocks), dim3(kProbeThreads), 0, nullptr, src, dst, nvec,
                           0x9E3779B9u, sink);
    }
}

template <int KIND>
inline void launch_probe_vw(int vec_bytes, bool nt, int blocks, const uint32_t *src, uint32_t *dst, size_t nbytes, uint32_t *sink) {
    switch (vec_bytes) {
        case 4: launch_probe<KIND, 1>(nt, blocks, src, dst, nbytes / 4, sink); break;
        case 8: launch_probe<KIND, 2>(nt, blocks, src, dst, nbytes / 8, sink); break;
        default: launch_probe<KIND, 4>(nt, blocks, src, dst, nbytes / 16, sink); break;
    }
}

}  // namespace mcg
