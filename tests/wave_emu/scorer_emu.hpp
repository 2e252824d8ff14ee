// Host stand-ins for the scorer's MFMA types (CPU wave emulator), included
// inside namespace mi.  The emulator runs the exact report path; the
// scorer's MFMA layers are checked on the GPU against a float32 reference.
struct alignas(16) bf16x8 {
    uint16_t v[8];
    uint16_t& operator[](int i) { return v[i]; }
};
struct f32x4 {
    float v[4];
    float& operator[](int i) { return v[i]; }
    const float& operator[](int i) const { return v[i]; }
};
inline bf16x8 pack8(float a0, float a1, float a2, float a3, float a4, float a5, float a6, float a7)
{
    return bf16x8{{bf16_bits(a0), bf16_bits(a1), bf16_bits(a2), bf16_bits(a3), bf16_bits(a4), bf16_bits(a5),
                   bf16_bits(a6), bf16_bits(a7)}};
}
inline f32x4 mfma16(bf16x8, bf16x8, f32x4) { __builtin_trap(); }
inline float __log2f(float x) { return __builtin_log2f(x); }
