// Candidate scorer MLP (no counterpart in the reference).
//
// A candidate is a read pair that passes report_thread's gate pre-check
// (dalign/filter.c:1740-1778); sixteen features of it (k_job_features) feed
// an MLP 16 -> 64 -> 64 -> 1 with ReLU hidden layers; bf16 weights and
// activations, fp32 accumulation.  The logit estimates whether the pair's
// exact run writes records.
//
// The three layers run on gfx950's v_mfma_f32_16x16x32_bf16 over a tile of
// 16 candidates (rows), one 16-column block at a time, so a wave holds a
// single 4-float accumulator.  Operand layout of that instruction, lane l of
// the wave:
//   A (16 x 32): row l & 15, k = 8*(l >> 4) + e, e = 0..7
//   B (32 x 16): column l & 15, k = 8*(l >> 4) + e
//   C (16 x 16): column l & 15, rows 4*(l >> 4) + v, v = 0..3
// Hidden activations go from the C layout of one layer to the A layout of the
// next through a 16 x 64 bf16 tile in LDS (2 KB per wave).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

constexpr int kFeat = 16;
constexpr int kHid = 64;
constexpr int kMlpRows = 16;                  // candidates per tile
constexpr int kMlpLdsHalves = kMlpRows * kHid;   // uint16 entries of one activation tile

// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t bf16_bits(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

#ifndef MI_CPU_WAVE_EMU
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float f) { return __builtin_bit_cast(__bf16, bf16_bits(f)); }

__device__ __forceinline__ bf16x8 pack8(float a0, float a1, float a2, float a3, float a4, float a5, float a6, float a7)
{
    bf16x8 r;
    r[0] = to_bf16(a0); r[1] = to_bf16(a1); r[2] = to_bf16(a2); r[3] = to_bf16(a3);
    r[4] = to_bf16(a4); r[5] = to_bf16(a5); r[6] = to_bf16(a6); r[7] = to_bf16(a7);
    return r;
}

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
#else
// host stand-ins of the CPU wave emulator (tests/wave_emu), which runs the
// exact report path only
#include "scorer_emu.hpp"
#endif

// Weights in MFMA operand order (one 16-byte B operand per lane and block).
struct MlpW {
    const bf16x8* w1;   // [4 column blocks][64 lanes]: W1 (16 x 64), k >= 16 zero
    const bf16x8* w2;   // [2 k-steps][4 column blocks][64 lanes]: W2 (64 x 64)
    const bf16x8* w3;   // [2 k-steps][64 lanes]: W3 (64 x 1) in column 0, zeros elsewhere
    const float* b1;    // [64]
    const float* b2;    // [64]
    float b3;
};

__device__ __forceinline__ void mlp_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// bias + ReLU of column block cb (C layout) into the LDS tile
__device__ __forceinline__ void mlp_store_block(f32x4 acc, int cb, const float* bias, uint16_t* lds)
{
    const int l = __lane_id();
    const int j = 16 * cb + (l & 15);
    const float b = bias[j];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float x = acc[v] + b;
        lds[(4 * (l >> 4) + v) * kHid + j] = bf16_bits(x > 0.f ? x : 0.f);
    }
}

// A operand of k-step t (columns 32t .. 32t+31) from the LDS tile
__device__ __forceinline__ bf16x8 mlp_load_a(const uint16_t* lds, int t)
{
    const int l = __lane_id();
    return *reinterpret_cast<const bf16x8*>(lds + (l & 15) * kHid + 32 * t + 8 * (l >> 4));
}

// Scores 16 candidates.  x: layer-1 A operand (row l & 15, features
// 8*(l>>4) .. +7; lanes 32..63 zero).  Returns the layer-3 accumulator: the
// logit of row r (before b3) sits in column 0, lane 16*(r >> 2), element
// r & 3.  lds: two 16 x 64 tiles.  All 64 lanes must be active.
__device__ __forceinline__ f32x4 mlp_tile(bf16x8 x, const MlpW& w, uint16_t* lds)
{
    const int l = __lane_id();
    const f32x4 zero = {};
    uint16_t* h1 = lds;
    uint16_t* h2 = lds + kMlpLdsHalves;
#pragma unroll
    for (int cb = 0; cb < 4; cb++) mlp_store_block(mfma16(x, w.w1[cb * 64 + l], zero), cb, w.b1, h1);
    mlp_wave_sync();
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        f32x4 g = zero;
#pragma unroll
        for (int t = 0; t < 2; t++) g = mfma16(mlp_load_a(h1, t), w.w2[(t * 4 + cb) * 64 + l], g);
        mlp_store_block(g, cb, w.b2, h2);
    }
    mlp_wave_sync();
    f32x4 o = zero;
#pragma unroll
    for (int t = 0; t < 2; t++) o = mfma16(mlp_load_a(h2, t), w.w3[t * 64 + l], o);
    mlp_wave_sync();
    return o;
}

}  // namespace mi
