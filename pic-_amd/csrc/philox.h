// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), written from the published round function.
// Shared by host setup code and device kernels so both draw identical streams.
//
// Stream layout used throughout the library:
//   key     = (seed_lo, seed_hi)
//   counter = (slot, step, tag, block)
// where `tag` packs species and purpose (see pic_stream_tag) and `block` advances
// by one per four 32-bit words consumed.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PIC_HD __host__ __device__ inline
#else
#define PIC_HD inline
#endif

struct philox4 { uint32_t v[4]; };

PIC_HD void pic_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

PIC_HD philox4 pic_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                 uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        pic_mulhilo32(M0, c0, &hi0, &lo0);
        pic_mulhilo32(M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    philox4 o; o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// Two 32-bit words -> double uniformly in the open interval (0,1): 52 random bits,
// centred in their bin, so the value is exact and never 0 or 1.
PIC_HD double pic_u01(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)a << 20) | (uint64_t)(b >> 12);   // 52 bits
    return ((double)m + 0.5) * (1.0 / 4503599627370496.0);    // 2^-52
}

// Sequential draws from one (slot, step, tag) stream.
struct pic_stream {
    uint32_t k0, k1, c0, c1, c2, blk;
    uint32_t buf[4];
    int used;
};

PIC_HD void pic_stream_init(pic_stream* s, uint64_t seed, uint32_t slot, uint32_t step, uint32_t tag) {
    s->k0 = (uint32_t)seed; s->k1 = (uint32_t)(seed >> 32);
    s->c0 = slot; s->c1 = step; s->c2 = tag; s->blk = 0; s->used = 4;
}

PIC_HD uint32_t pic_stream_u32(pic_stream* s) {
    if (s->used == 4) {
        philox4 o = pic_philox4x32_10(s->c0, s->c1, s->c2, s->blk, s->k0, s->k1);
        s->buf[0] = o.v[0]; s->buf[1] = o.v[1]; s->buf[2] = o.v[2]; s->buf[3] = o.v[3];
        s->blk++; s->used = 0;
    }
    return s->buf[s->used++];
}

PIC_HD double pic_stream_u01(pic_stream* s) {
    uint32_t a = pic_stream_u32(s);
    uint32_t b = pic_stream_u32(s);
    return pic_u01(a, b);
}

// Purposes packed into the counter's third word together with the species index.
enum {
    PIC_RNG_LOAD = 1,      // initial Maxwellian load
    PIC_RNG_EMIT = 2,      // left-wall kappa re-emission inside the push
    PIC_RNG_SELECT = 3,    // null-collision selection (one word per slot)
    PIC_RNG_COLLIDE = 4,   // collision kinematics of a selected particle
    PIC_RNG_INJECT = 5,    // wall injection
    PIC_RNG_LOADV = 6      // kappa / polytropic velocities of the initial load
};

PIC_HD uint32_t pic_stream_tag(int species, int purpose) {
    return ((uint32_t)purpose << 16) | (uint32_t)(species & 0xFFFF);
}
