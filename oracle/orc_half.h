/*
 * ORACLE (test infrastructure only).  BF16 and IEEE binary16 storage conversions,
 * written from the formats' definitions: round to nearest, ties to even; binary16
 * keeps subnormals and rounds values of magnitude >= 65520 to infinity; NaN stays NaN
 * (quiet).  Used to restate the reduced-precision storage form of psi and of the ADE
 * P/J state, whose arithmetic is FP32.
 */
static inline uint32_t orc_f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float orc_u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

static inline float orc_bf16_to_f32(uint16_t h) { return orc_u2f((uint32_t)h << 16); }

static inline uint16_t orc_f32_to_bf16(float x) {
    uint32_t u = orc_f2u(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static inline float orc_f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {                                   /* zero or subnormal: m * 2^-24 */
        const float v = (float)m * 5.9604644775390625e-8f;
        return orc_u2f(orc_f2u(v) | sign);
    }
    if (e == 31) return orc_u2f(sign | 0x7f800000u | (m << 13));
    return orc_u2f(sign | ((e + 112u) << 23) | (m << 13));
}

static inline uint16_t orc_f32_to_f16(float x) {
    uint32_t u = orc_f2u(x);
    const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((u >> 13) & 0x3ffu));
    if (u >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);        /* >= 65520 -> inf */
    if (u >= 0x38800000u) {                                          /* normal (>= 2^-14) */
        u += 0xfffu + ((u >> 13) & 1u);
        return (uint16_t)(sign | ((u - 0x38000000u) >> 13));
    }
    /* subnormal range: round |x| / 2^-24 to an integer, ties to even (exact scaling) */
    const float q = orc_u2f(u) * 16777216.0f;
    return (uint16_t)(sign | (uint16_t)rintf(q));
}

/* storage load/store of one element of kind MP_F32, MP_F64, MP_BF16 or MP_F16; reduced
 * kinds hold value*sc and are read back as stored*isc, both products in FP32 */
static inline double orc_sld(const void *p, int64_t o, int kind, float isc) {
    switch (kind) {
    case MP_F32: return ((const float *)p)[o];
    case MP_F64: return ((const double *)p)[o];
    case MP_BF16: return orc_bf16_to_f32(((const uint16_t *)p)[o]) * isc;
    default: return orc_f16_to_f32(((const uint16_t *)p)[o]) * isc;
    }
}
static inline void orc_sst(void *p, int64_t o, int kind, double v, float sc) {
    switch (kind) {
    case MP_F32: ((float *)p)[o] = (float)v; break;
    case MP_F64: ((double *)p)[o] = v; break;
    case MP_BF16: ((uint16_t *)p)[o] = orc_f32_to_bf16((float)v * sc); break;
    default: ((uint16_t *)p)[o] = orc_f32_to_f16((float)v * sc); break;
    }
}
