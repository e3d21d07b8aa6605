// rf_philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), used as the
// library's device phase stream when the caller supplies no draws.  Host and
// device callable so the host test suite can check known-answer vectors.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define RF_HD __host__ __device__ inline
#else
#define RF_HD inline
#endif

struct rf_u32x4 { uint32_t v[4]; };

RF_HD void rf_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

// One Philox S-box round: two 32x32->64 multiplies, xor with the key.
RF_HD rf_u32x4 rf_philox_round(rf_u32x4 c, uint32_t k0, uint32_t k1) {
    uint32_t hi0, lo0, hi1, lo1;
    rf_mulhilo32(0xD2511F53u, c.v[0], &hi0, &lo0);
    rf_mulhilo32(0xCD9E8D57u, c.v[2], &hi1, &lo1);
    rf_u32x4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1;
    o.v[3] = lo0;
    return o;
}

RF_HD rf_u32x4 rf_philox4x32_10(rf_u32x4 c, uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }   // Weyl key schedule
        c = rf_philox_round(c, k0, k1);
    }
    return c;
}

// Element e of stream (seed, tag): U[0,1) with a 53-bit mantissa.
RF_HD double rf_uniform(uint64_t seed, uint32_t tag, uint64_t e) {
    rf_u32x4 c;
    c.v[0] = (uint32_t)e;
    c.v[1] = (uint32_t)(e >> 32);
    c.v[2] = tag;
    c.v[3] = 0u;
    const rf_u32x4 o = rf_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t m = ((uint64_t)o.v[0] << 21) | (uint64_t)(o.v[1] >> 11);
    return (double)m * (1.0 / 9007199254740992.0);
}
