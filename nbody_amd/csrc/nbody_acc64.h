// Device code of the FP64-accumulation plans (compensated = 2): the
// symmetric block pairs and the LDS-tiled one-sided sweep with every term
// added in FP64.
//
// Each term's inputs are the FP32 values of the default kernels: d = x_j - x_i,
// w = m (inv inv) inv or m inv (inv inv) with inv = v_rsq_f32(|d|^2 + eps2).
// Their product is exact in FP64 (24 x 24 bits), and a sum of such products
// in FP64 loses about 2^-53 of the running sum per term, so the force keeps
// only the FP32 rounding of each term's own inputs, which does not grow with
// cancellation between terms. The FP32 plans round every partial sum to 24
// bits, which is what costs digits on a body whose force is the small
// remainder of large cancelling terms. Partial sums (slots, slices) stay in
// FP64 until the one final rounding to the FP32 force.
#pragma once

#include "nbody_device.h"
#include "nbody_sym.h"

namespace nbody {

template <int I>
struct Lanes64 {
    double x[I], y[I], z[I];
};

template <int I>
__device__ __forceinline__ void zero(Lanes64<I>& a)
{
#pragma unroll
    for (int u = 0; u < I; ++u) a.x[u] = a.y[u] = a.z[u] = 0.0;
}

typedef SlotMapT<double> SlotMap64;

// wave_ror1 of a double: both 32-bit halves move one lane.
__device__ __forceinline__ double wave_ror1_f64(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x13C, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x13C, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// The tiles of block pair (P, Q) against this lane's A bodies of P (see
// sym_pair64). With EQ every body of both blocks has mass m0: the term
// weights leave it out (inv^3 on the i side, -inv^3 on the j side: one
// conversion to FP64 instead of three, no mass products, the masses do not
// rotate) and the j-side partials are multiplied by m0 at the store; for
// m0 = 1 the terms are the general path's bits.
template <int A, bool NEWTON, bool EQ>
__device__ __forceinline__ void sym_tiles64(const float4* __restrict__ pm, int n, float eps2, const SlotMap64& slots,
                                            int P, int Q, const Lanes<A>& me, const float (&mi)[A], double m0,
                                            Lanes64<A>& acc)
{
    constexpr int B = 64 * A;
    const int lane = threadIdx.x;
    for (int tile = 0; tile < A; ++tile) {
        const int jbase = Q * B + tile * 64;  // wave-uniform
        if (jbase >= n) break;
        const float4 b = lane < n - jbase ? pm[jbase + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
        float xj = b.x, yj = b.y, zj = b.z, mj = b.w;
        double rx = 0.0, ry = 0.0, rz = 0.0;
        for (int k = 0; k < 64; ++k) {
#pragma unroll
            for (int u = 0; u < A; ++u) {
                const float dx = xj - me.x[u];
                const float dy = yj - me.y[u];
                const float dz = zj - me.z[u];
                float d2 = __builtin_fmaf(dx, dx, eps2);
                d2 = __builtin_fmaf(dy, dy, d2);
                d2 = __builtin_fmaf(dz, dz, d2);
                const float inv = rsqrt_f32<NEWTON>(d2);
                const float inv3 = (inv * inv) * inv;
                const double DX = dx, DY = dy, DZ = dz;
                if constexpr (EQ) {
                    const double w = (double)inv3;
                    acc.x[u] = __builtin_fma(DX, w, acc.x[u]);
                    acc.y[u] = __builtin_fma(DY, w, acc.y[u]);
                    acc.z[u] = __builtin_fma(DZ, w, acc.z[u]);
                    rx = __builtin_fma(DX, -w, rx);
                    ry = __builtin_fma(DY, -w, ry);
                    rz = __builtin_fma(DZ, -w, rz);
                } else {
                    const double si = (double)(mj * inv3);
                    const double sj = (double)(-mi[u] * inv3);
                    acc.x[u] = __builtin_fma(DX, si, acc.x[u]);
                    acc.y[u] = __builtin_fma(DY, si, acc.y[u]);
                    acc.z[u] = __builtin_fma(DZ, si, acc.z[u]);
                    rx = __builtin_fma(DX, sj, rx);
                    ry = __builtin_fma(DY, sj, ry);
                    rz = __builtin_fma(DZ, sj, rz);
                }
            }
            xj = wave_ror1(xj);
            yj = wave_ror1(yj);
            zj = wave_ror1(zj);
            if constexpr (!EQ) mj = wave_ror1(mj);
            rx = wave_ror1_f64(rx);
            ry = wave_ror1_f64(ry);
            rz = wave_ror1_f64(rz);
        }
        double* __restrict__ jdst = slots.j_side(P, Q, B) + tile * 64 + lane;
        jdst[0] = EQ ? rx * m0 : rx;
        jdst[slots.bplane] = EQ ? ry * m0 : ry;
        jdst[2 * slots.bplane] = EQ ? rz * m0 : rz;
    }
}

// Block pair (P, Q), P < Q, on one wave with A bodies of block P per lane
// (B = 64 A): the rotation of nbody_sym.h (the j tile's position, mass and
// FP64 reaction move one lane per step), each term accumulated in FP64. Every
// tile is visited by this wave alone, so its reactions go straight from the
// lanes to their slots. Tiles past the end of a ragged last block are
// skipped: their bodies are padding, whose slots no reduction reads. When
// every body of both blocks has the mass of body P*B (a wave vote; padding
// has mass 0, so a ragged block never qualifies) the mass-free terms run.
template <int A, bool NEWTON>
__device__ __forceinline__ void sym_pair64(const float4* __restrict__ pm, int n, float eps2, const SlotMap64& slots,
                                           int P, int Q)
{
    constexpr int B = 64 * A;
    const int lane = threadIdx.x;
    const float m0 = pm[P * B].w;
    Lanes<A> me;
    float mi[A];
    bool same = true;
#pragma unroll
    for (int u = 0; u < A; ++u) {
        const float4 q = pm[P * B + u * 64 + lane];
        me.x[u] = q.x;
        me.y[u] = q.y;
        me.z[u] = q.z;
        mi[u] = q.w;
        same &= q.w == m0;
    }
#pragma unroll
    for (int u = 0; u < A; ++u) {
        const int j = Q * B + u * 64 + lane;
        same &= (j < n) & (pm[min(j, n - 1)].w == m0);
    }
    Lanes64<A> acc;
    zero<A>(acc);
    const bool eq = __builtin_amdgcn_ballot_w64(!same) == 0;
    if (eq) sym_tiles64<A, NEWTON, true>(pm, n, eps2, slots, P, Q, me, mi, (double)m0, acc);
    else sym_tiles64<A, NEWTON, false>(pm, n, eps2, slots, P, Q, me, mi, 1.0, acc);
    const double scale = eq ? (double)m0 : 1.0;
    double* __restrict__ idst = slots.i_side(P, Q, B) + lane;
    const size_t iplane = slots.i_plane(P);
#pragma unroll
    for (int u = 0; u < A; ++u) {
        idst[u * 64] = acc.x[u] * scale;
        idst[iplane + u * 64] = acc.y[u] * scale;
        idst[2 * iplane + u * 64] = acc.z[u] * scale;
    }
}

// Body i (this lane's) against bodies [j0, j0 + len), one-sided, in j order,
// every term added in FP64 into (ax, ay, az). A lane whose i is past n
// computes on a zero body.
template <bool NEWTON>
__device__ __forceinline__ void row64(const float4* __restrict__ pm, int n, float eps2, int i, int j0, int len,
                                      double& ax, double& ay, double& az)
{
    const float4 q = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    ax = ay = az = 0.0;
    for (int j = 0; j < len; ++j) {
        const float4 b = pm[j0 + j];
        const float dx = b.x - q.x;
        const float dy = b.y - q.y;
        const float dz = b.z - q.z;
        float d2 = __builtin_fmaf(dx, dx, eps2);
        d2 = __builtin_fmaf(dy, dy, d2);
        d2 = __builtin_fmaf(dz, dz, d2);
        const float inv = rsqrt_f32<NEWTON>(d2);
        const double w = (double)(b.w * ((inv * inv) * inv));
        ax = __builtin_fma((double)dx, w, ax);
        ay = __builtin_fma((double)dy, w, ay);
        az = __builtin_fma((double)dz, w, az);
    }
}

// Piece k (of A) of block P's diagonal: 64 bodies P*B + 64k + lane against
// the bodies of block P, one-sided, in j order, into slot P. Lanes past n
// compute on a zero body and store into padding that no reduction reads.
template <int B, bool NEWTON>
__device__ __forceinline__ void sym_diag_piece64(const float4* __restrict__ pm, int n, float eps2,
                                                 const SlotMap64& slots, int P, int k)
{
    if (P * B + k * 64 >= n) return;  // a piece made only of padding (wave-uniform)
    const int i = P * B + k * 64 + (int)threadIdx.x;
    const int j0 = P * B;
    double ax, ay, az;
    row64<NEWTON>(pm, n, eps2, i, j0, min(n - j0, B), ax, ay, az);
    double* __restrict__ dst = slots.j_side(P, P, B) + k * 64 + threadIdx.x;
    dst[0] = ax;
    dst[slots.bplane] = ay;
    dst[2 * slots.bplane] = az;
}

}  // namespace nbody
