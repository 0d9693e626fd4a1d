// mavi_philox.h -- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), for
// the device kernels and for host-side draws of the same streams.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// One round of Philox4x32 (philox_normal adds Box-Muller on two 53-bit uniforms).
__host__ __device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                                      uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint32_t hi0 = (uint32_t)(((uint64_t)M0 * c0) >> 32), lo0 = M0 * c0;
    uint32_t hi1 = (uint32_t)(((uint64_t)M1 * c2) >> 32), lo1 = M1 * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
}

// Two uniforms in [0, 1) with 53 random bits each from Philox4x32-10 keyed
// by (seed, step, id); a counter word different from philox_normal's keeps
// the streams apart.
__host__ __device__ inline void philox_uniform2(uint64_t seed, uint64_t step, uint32_t id, double *u, double *w) {
    uint32_t c0 = (uint32_t)step, c1 = (uint32_t)(step >> 32), c2 = id, c3 = 0x72746d62u;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    *u = (double)((((uint64_t)c0 << 21) ^ ((uint64_t)c1 >> 11)) & ((1ull << 53) - 1)) * 0x1p-53;
    *w = (double)((((uint64_t)c2 << 21) ^ ((uint64_t)c3 >> 11)) & ((1ull << 53) - 1)) * 0x1p-53;
}

