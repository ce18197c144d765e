// Device-side building blocks shared by every force kernel (gfx950, wave64).
//
// The interaction is the reference's bodyBodyInteraction
// (src/bodybodyInteraction.cuh:40-65): f_i += m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^(3/2),
// without m_i and without G. It is written here as 12 VALU ops + 1 v_rsq_f32:
//   3 v_sub (d = x_j - x_i), 3 v_fma (d2 = eps2 + dx^2 + dy^2 + dz^2),
//   v_rsq_f32, 3 v_mul (s = (m_j inv) (inv inv)), 3 v_fma (a += d s).
// Differences from the reference's FP32 rounding sequence, all within rounding:
// eps2 enters the FMA chain first, the sums are fused, and the product is
// grouped (m inv)(inv inv). v_rsq_f32 measures <= 0.86 ulp on gfx950, so the
// optional Newton step (NEWTON) is off by default.
#pragma once

#include <hip/hip_runtime.h>

namespace nbody {

// Knuth's TwoSum: returns s = fl(a + b) and sets err so that s + err == a + b
// exactly (round-to-nearest, no overflow). Six adds; no reassociation is
// allowed in this translation unit (no fast-math).
__device__ __forceinline__ float two_sum(float a, float b, float& err)
{
    const float s = a + b;
    const float bv = s - a;
    const float av = s - bv;
    err = (a - av) + (b - bv);
    return s;
}

template <bool NEWTON>
__device__ __forceinline__ float rsqrt_f32(float d2)
{
    float y = __builtin_amdgcn_rsqf(d2);
    if constexpr (NEWTON) {
        // one Newton-Raphson step in residual form: y += (y/2)(1 - d2 y^2)
        const float e = __builtin_fmaf(-(d2 * y), y, 1.0f);
        y = __builtin_fmaf(0.5f * y, e, y);
    }
    return y;
}

template <bool NEWTON>
__device__ __forceinline__ void body_body(float xj, float yj, float zj, float mj,
                                          float xi, float yi, float zi, float eps2,
                                          float& ax, float& ay, float& az)
{
    const float dx = xj - xi;
    const float dy = yj - yi;
    const float dz = zj - zi;
    float d2 = __builtin_fmaf(dx, dx, eps2);
    d2 = __builtin_fmaf(dy, dy, d2);
    d2 = __builtin_fmaf(dz, dz, d2);
    const float inv = rsqrt_f32<NEWTON>(d2);
    const float s = (mj * inv) * (inv * inv);
    ax = __builtin_fmaf(dx, s, ax);
    ay = __builtin_fmaf(dy, s, ay);
    az = __builtin_fmaf(dz, s, az);
}

// The same interaction as one fixed dependent chain in assembly: gfx950
// issues a chain whose instructions consume each other's results back to back
// at a measurably higher rate than the compiler's interleaving of independent
// interactions. Same operations and rounding as body_body<false> (bit-
// identical); one s_nop after v_rsq_f32 is the wait state gfx950 requires
// before a VALU reads a transcendental result. JC is the register class of
// the j body: "v" (from LDS or a DPP tile) or "s" (from scalar loads).
#define NBODY_BODY_CHAIN(NAME, JC)                                                                     \
    __device__ __forceinline__ void NAME(float xj, float yj, float zj, float mj, float xi, float yi,   \
                                         float zi, float eps2, float& ax, float& ay, float& az)        \
    {                                                                                                  \
        float dx, dy, dz, s, t;                                                                        \
        asm volatile("v_sub_f32 %[dx], %[xj], %[xi]\n"                                                 \
                     "v_sub_f32 %[dy], %[yj], %[yi]\n"                                                 \
                     "v_sub_f32 %[dz], %[zj], %[zi]\n"                                                 \
                     "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                          \
                     "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                 \
                     "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                 \
                     "v_rsq_f32 %[s], %[s]\n"                                                          \
                     "s_nop 0\n"                                                                       \
                     "v_mul_f32 %[t], %[mj], %[s]\n"                                                   \
                     "v_mul_f32 %[s], %[s], %[s]\n"                                                    \
                     "v_mul_f32 %[s], %[t], %[s]\n"                                                    \
                     "v_fmac_f32 %[ax], %[dx], %[s]\n"                                                 \
                     "v_fmac_f32 %[ay], %[dy], %[s]\n"                                                 \
                     "v_fmac_f32 %[az], %[dz], %[s]\n"                                                 \
                     : [ax] "+v"(ax), [ay] "+v"(ay), [az] "+v"(az), [dx] "=&v"(dx), [dy] "=&v"(dy),     \
                       [dz] "=&v"(dz), [s] "=&v"(s), [t] "=&v"(t)                                      \
                     : [xj] JC(xj), [yj] JC(yj), [zj] JC(zj), [mj] JC(mj), [xi] "v"(xi), [yi] "v"(yi),  \
                       [zi] "v"(zi), [eps] "s"(eps2));                                                 \
    }
NBODY_BODY_CHAIN(body_body_chain_v, "v")
NBODY_BODY_CHAIN(body_body_chain_s, "s")
#undef NBODY_BODY_CHAIN

// The chain without the mass, for j ranges whose bodies all have one mass m0
// (the caller multiplies the sums by m0 once): inv^3 = inv * (inv * inv),
// the product the chain above forms for m = 1, so with m0 = 1 (or any power
// of two) the sums have the same bits. 11 VALU + 1 v_rsq_f32. The mass is an
// operand the chain does not read: with it in use, the j record stays one
// 16-byte load (gfx950 has no 12-byte scalar load, and the scalar-broadcast
// kernel would otherwise get vector loads for x, y, z).
#define NBODY_BODY_CHAIN_EQ(NAME, JC)                                                                  \
    __device__ __forceinline__ void NAME(float xj, float yj, float zj, float mj, float xi, float yi,   \
                                         float zi, float eps2, float& ax, float& ay, float& az)        \
    {                                                                                                  \
        float dx, dy, dz, s, t;                                                                        \
        asm volatile("v_sub_f32 %[dx], %[xj], %[xi]\n"                                                 \
                     "v_sub_f32 %[dy], %[yj], %[yi]\n"                                                 \
                     "v_sub_f32 %[dz], %[zj], %[zi]\n"                                                 \
                     "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                          \
                     "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                 \
                     "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                 \
                     "v_rsq_f32 %[s], %[s]\n"                                                          \
                     "s_nop 0\n"                                                                       \
                     "v_mul_f32 %[t], %[s], %[s]\n"                                                    \
                     "v_mul_f32 %[s], %[s], %[t]\n"                                                    \
                     "v_fmac_f32 %[ax], %[dx], %[s]\n"                                                 \
                     "v_fmac_f32 %[ay], %[dy], %[s]\n"                                                 \
                     "v_fmac_f32 %[az], %[dz], %[s]\n"                                                 \
                     : [ax] "+v"(ax), [ay] "+v"(ay), [az] "+v"(az), [dx] "=&v"(dx), [dy] "=&v"(dy),     \
                       [dz] "=&v"(dz), [s] "=&v"(s), [t] "=&v"(t)                                      \
                     : [xj] JC(xj), [yj] JC(yj), [zj] JC(zj), [mj] JC(mj), [xi] "v"(xi), [yi] "v"(yi),  \
                       [zi] "v"(zi), [eps] "s"(eps2));                                                 \
    }
NBODY_BODY_CHAIN_EQ(body_body_chain_eq_v, "v")
NBODY_BODY_CHAIN_EQ(body_body_chain_eq_s, "s")
#undef NBODY_BODY_CHAIN_EQ

// body_body for the kernels: the assembly chain when no Newton step is asked
// for, with the j body in VGPRs (SGPR_J false) or SGPRs (SGPR_J true). EQ
// leaves the mass out (the mass-free chain; with a Newton step, m = 1).
template <bool NEWTON, bool SGPR_J, bool EQ = false>
__device__ __forceinline__ void interact(float xj, float yj, float zj, float mj, float xi, float yi, float zi,
                                         float eps2, float& ax, float& ay, float& az)
{
    if constexpr (NEWTON) body_body<true>(xj, yj, zj, EQ ? 1.0f : mj, xi, yi, zi, eps2, ax, ay, az);
    else if constexpr (EQ && SGPR_J) body_body_chain_eq_s(xj, yj, zj, mj, xi, yi, zi, eps2, ax, ay, az);
    else if constexpr (EQ) body_body_chain_eq_v(xj, yj, zj, mj, xi, yi, zi, eps2, ax, ay, az);
    else if constexpr (SGPR_J) body_body_chain_s(xj, yj, zj, mj, xi, yi, zi, eps2, ax, ay, az);
    else body_body_chain_v(xj, yj, zj, mj, xi, yi, zi, eps2, ax, ay, az);
}

// Per-lane state of I i bodies (or their accumulators).
template <int I>
struct Lanes {
    float x[I], y[I], z[I];
};

template <int I>
__device__ __forceinline__ void zero(Lanes<I>& a)
{
#pragma unroll
    for (int u = 0; u < I; ++u) a.x[u] = a.y[u] = a.z[u] = 0.f;
}

// hi += part with TwoSum, residue into lo.
template <int I>
__device__ __forceinline__ void fold(Lanes<I>& hi, Lanes<I>& lo, const Lanes<I>& part)
{
#pragma unroll
    for (int u = 0; u < I; ++u) {
        float e;
        hi.x[u] = two_sum(hi.x[u], part.x[u], e);
        lo.x[u] += e;
        hi.y[u] = two_sum(hi.y[u], part.y[u], e);
        lo.y[u] += e;
        hi.z[u] = two_sum(hi.z[u], part.z[u], e);
        lo.z[u] += e;
    }
}

// j bodies read with wave-uniform addresses: the compiler emits s_load_dwordx16
// (four bodies per load) and the VALU ops take x_j, y_j, z_j, m_j as SGPR
// operands, so no instruction is spent on the broadcast; on gfx950 a VALU op
// that reads an SGPR issues at about 60% of the all-VGPR rate, though.
// One sweep over j bodies in groups of G: the G loads are issued before the
// G x I interaction chains that consume them (the chains are assembly the
// compiler neither reorders nor unrolls around), then the remainder.
template <int I, bool NEWTON, bool SGPR_J, int G, bool EQ = false, class Load>
__device__ __forceinline__ void sweep_grouped(const Load& load, int len, float eps2, const Lanes<I>& me,
                                              Lanes<I>& acc)
{
    int j = 0;
    for (; j + G <= len; j += G) {
        float4 b[G];
#pragma unroll
        for (int q = 0; q < G; ++q) b[q] = load(j + q);
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int u = 0; u < I; ++u)
                interact<NEWTON, SGPR_J, EQ>(b[q].x, b[q].y, b[q].z, b[q].w, me.x[u], me.y[u], me.z[u], eps2,
                                             acc.x[u], acc.y[u], acc.z[u]);
    }
    for (; j < len; ++j) {
        const float4 b = load(j);
#pragma unroll
        for (int u = 0; u < I; ++u)
            interact<NEWTON, SGPR_J, EQ>(b.x, b.y, b.z, b.w, me.x[u], me.y[u], me.z[u], eps2, acc.x[u], acc.y[u],
                                         acc.z[u]);
    }
}

// j bodies read with wave-uniform addresses: the compiler emits s_load_dwordx16
// (four bodies per load) and the VALU ops take x_j, y_j, z_j, m_j as SGPR
// operands, so no instruction is spent on the broadcast.
template <int I, bool NEWTON, bool EQ = false>
__device__ __forceinline__ void sweep_sgpr(const float4* __restrict__ src, int len, float eps2,
                                           const Lanes<I>& me, Lanes<I>& acc)
{
    sweep_grouped<I, NEWTON, true, 8, EQ>([src](int k) { return src[k]; }, len, eps2, me, acc);
}

// Read-only view of body records for kernels whose body pointers are not
// kernel arguments (batched tasks carry them in memory): the constant address
// space tells the compiler no lane writes them during the launch, so the
// wave-uniform reads stay scalar loads (s_load) into SGPR operands.
typedef const __attribute__((address_space(4))) float const_body_float;

template <int I, bool NEWTON, bool EQ = false>
__device__ __forceinline__ void sweep_const(const float4* src4, int len, float eps2, const Lanes<I>& me,
                                            Lanes<I>& acc)
{
    const const_body_float* src = (const const_body_float*)(const void*)src4;
    sweep_grouped<I, NEWTON, true, 8, EQ>(
        [src](int k) { return make_float4(src[4 * k + 0], src[4 * k + 1], src[4 * k + 2], src[4 * k + 3]); }, len,
        eps2, me, acc);
}

}  // namespace nbody
