// Force, integrate and FP64 diagnostic kernels for gfx950 (MI355X), wave64.
//
// Decomposition shared by the FP32 force kernels (every variant):
//   * block = 256 lanes = 4 waves; lane t of block b owns i bodies
//     m = b*256*I + u*256 + t (u < I), so each wave's i loads and stores are
//     contiguous float4 runs;
//   * blockIdx.y selects a j slice [y*jslice, min(n, (y+1)*jslice));
//   * the variants differ only in how the j bodies reach the lanes: one
//     wave-uniform body at a time (scalar loads into SGPRs, an LDS
//     broadcast), or a 64-body tile that DPP moves one lane per step;
//   * slice partials (float4, plus a float4 of TwoSum residues when
//     compensated) are summed by k_reduce in slice order, so a result depends
//     on the plan only, never on scheduling.
// With one slice the kernel writes force directly. Force records are written
// x, y, z only; w is left alone (the contract of src/nbody.cu:351-371).
#include "nbody_acc64.h"
#include "nbody_device.h"
#include "nbody_kernels.h"
#include "nbody_sym.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <utility>

namespace nbody {

namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 128;  // j bodies between TwoSum folds (multiple of 16)

template <int I>
__device__ __forceinline__ void load_i(const ForceParams& p, Lanes<I>& me)
{
    const int mbase = blockIdx.x * (kBlock * I) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < I; ++u) {
        const int m = min(mbase + u * kBlock, p.ni - 1);
        const float4 q = p.pm[p.i0 + m];
        me.x[u] = q.x;
        me.y[u] = q.y;
        me.z[u] = q.z;
    }
}

template <int I>
__device__ __forceinline__ void store(const ForceParams& p, const Lanes<I>& hi, const Lanes<I>* lo)
{
    const int mbase = blockIdx.x * (kBlock * I) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < I; ++u) {
        const int m = mbase + u * kBlock;
        if (m >= p.ni) continue;
        if (p.hi == nullptr) {
            float* f = p.force + 4 * (size_t)m;
            if (lo) {
                f[0] = hi.x[u] + lo->x[u];
                f[1] = hi.y[u] + lo->y[u];
                f[2] = hi.z[u] + lo->z[u];
            } else {
                f[0] = hi.x[u];
                f[1] = hi.y[u];
                f[2] = hi.z[u];
            }
        } else {
            const size_t k = (size_t)blockIdx.y * p.ni + m;
            p.hi[k] = make_float4(hi.x[u], hi.y[u], hi.z[u], 0.f);
            if (lo) p.lo[k] = make_float4(lo->x[u], lo->y[u], lo->z[u], 0.f);
        }
    }
}

// True (on every lane) when every body of the slice src[0, len) has the mass
// of its first, m0 (uniform initial conditions: unit masses, m = 1/N). The
// one-sided kernels then run the mass-free chain (11 VALU + 1 v_rsq_f32 per
// interaction instead of 12) and multiply the slice's sums by m0 once; for
// m0 a power of two that gives the same bits as the general chain, otherwise
// it differs by the rounding of that one product. Each wave votes on its own
// (a ballot, no barrier): a barrier would count as a possible store to the
// bodies, and the SGPR kernel's wave-uniform j loads after it would no longer
// be scalar loads.
__device__ __forceinline__ bool uniform_mass(const float4* __restrict__ src, int len, float& m0)
{
    m0 = src[0].w;
    bool ok = true;
    for (int k = threadIdx.x & 63; k < len; k += 64) ok &= src[k].w == m0;
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

template <int I>
__device__ __forceinline__ void scale(Lanes<I>& a, float m)
{
#pragma unroll
    for (int u = 0; u < I; ++u) {
        a.x[u] *= m;
        a.y[u] *= m;
        a.z[u] *= m;
    }
}

// A slice's sums into hi through sweep(c, len, acc), which adds the terms of
// bodies [c, c + len) of the slice: one plain sum, or (COMP) plain sums of
// kChunk terms folded into the TwoSum pair (hi, lo).
template <int I, bool COMP, class Sweep>
__device__ __forceinline__ void slice_sums(const Sweep& sweep, int len, Lanes<I>& hi, Lanes<I>& lo)
{
    if constexpr (!COMP) {
        sweep(0, len, hi);
    } else {
        for (int c = 0; c < len; c += kChunk) {
            Lanes<I> part;
            zero<I>(part);
            sweep(c, min(kChunk, len - c), part);
            fold<I>(hi, lo, part);
        }
    }
}

// slice_sums with the mass-free chain when the slice's masses are all m0
// (then the sums are multiplied by m0), else the general chain; SWEEP<EQ>
// is the kernel's sweep. The branch is taken once per slice, outside the
// loops, so each path keeps its own straight loop.
template <int I, bool COMP, class SweepEq, class SweepGen>
__device__ __forceinline__ void slice_sums_mass(const float4* __restrict__ src, int len, const SweepEq& eq_sweep,
                                                const SweepGen& gen_sweep, Lanes<I>& hi, Lanes<I>& lo)
{
    float m0;
    if (uniform_mass(src, len, m0)) {
        slice_sums<I, COMP>(eq_sweep, len, hi, lo);
        scale<I>(hi, m0);
        if constexpr (COMP) scale<I>(lo, m0);
    } else {
        slice_sums<I, COMP>(gen_sweep, len, hi, lo);
    }
}

template <int I, bool COMP, bool NEWTON>
__global__ __launch_bounds__(kBlock) void k_force_sgpr(ForceParams p)
{
    Lanes<I> me, hi, lo;
    load_i<I>(p, me);
    zero<I>(hi);
    zero<I>(lo);
    const int j0 = blockIdx.y * p.jslice;
    const int len = min(p.n - j0, p.jslice);
    const float4* __restrict__ src = p.pm + j0;
    const float eps2 = p.eps2;
    // j bodies through the constant address space: with the mass vote's
    // vector loads of the slice in the kernel, the compiler keeps only one of
    // the two sweeps' wave-uniform loads scalar otherwise
    slice_sums_mass<I, COMP>(
        src, len, [&](int c, int l, Lanes<I>& a) { sweep_const<I, NEWTON, true>(src + c, l, eps2, me, a); },
        [&](int c, int l, Lanes<I>& a) { sweep_const<I, NEWTON>(src + c, l, eps2, me, a); }, hi, lo);
    store<I>(p, hi, COMP ? &lo : nullptr);
}

// The reference's first rung (src/nbody_GPU_AOS.cu:42-75): one i body per
// lane, the whole j range, no slices.
template <bool NEWTON>
__global__ __launch_bounds__(kBlock) void k_force_aos(ForceParams p)
{
    Lanes<1> me, acc;
    load_i<1>(p, me);
    zero<1>(acc);
    for (int j = 0; j < p.n; ++j) {
        const float4 b = p.pm[j];
        body_body<NEWTON>(b.x, b.y, b.z, b.w, me.x[0], me.y[0], me.z[0], p.eps2, acc.x[0], acc.y[0],
                          acc.z[0]);
    }
    store<1>(p, acc, nullptr);
}

// j tiles of 256 bodies staged in LDS by the whole block (one coalesced float4
// per lane), double-buffered so one barrier per tile suffices; each wave then
// reads body k with a uniform address (a broadcast ds_read_b128).
template <int I, bool COMP, bool NEWTON>
__global__ __launch_bounds__(kBlock) void k_force_lds(ForceParams p)
{
    __shared__ float4 tile[2][kBlock];
    Lanes<I> me, hi, lo;
    load_i<I>(p, me);
    zero<I>(hi);
    zero<I>(lo);
    const int j0 = blockIdx.y * p.jslice;
    const int len = min(p.n - j0, p.jslice);
    const float4* __restrict__ src = p.pm + j0;
    float m0;
    const bool eq = uniform_mass(src, len, m0);
    int buf = 0;
    for (int t = 0; t < len; t += kBlock, buf ^= 1) {
        const int k = t + threadIdx.x;
        tile[buf][threadIdx.x] = k < len ? src[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const int cnt = min(kBlock, len - t);
        Lanes<I> part;
        zero<I>(part);
        Lanes<I>& acc = COMP ? part : hi;
        const float4* t4 = tile[buf];
        if (eq) sweep_grouped<I, NEWTON, false, 4, true>([t4](int q) { return t4[q]; }, cnt, p.eps2, me, acc);
        else sweep_grouped<I, NEWTON, false, 4>([t4](int q) { return t4[q]; }, cnt, p.eps2, me, acc);
        if constexpr (COMP) fold<I>(hi, lo, part);
    }
    if (eq) {
        scale<I>(hi, m0);
        if constexpr (COMP) scale<I>(lo, m0);
    }
    store<I>(p, hi, COMP ? &lo : nullptr);
}

// k_force_lds with every term added in FP64 (the accurate plan): the same
// tiles and term inputs, FP64 accumulators, FP64 slice partials
// hi64[slice][component][ni] (or the force, rounded once, with one slice).
template <int I, bool NEWTON>
__global__ __launch_bounds__(kBlock) void k_force_lds64(ForceParams p)
{
    __shared__ float4 tile[2][kBlock];
    Lanes<I> me;
    Lanes64<I> acc;
    load_i<I>(p, me);
    zero<I>(acc);
    const int j0 = blockIdx.y * p.jslice;
    const int len = min(p.n - j0, p.jslice);
    const float4* __restrict__ src = p.pm + j0;
    int buf = 0;
    for (int t = 0; t < len; t += kBlock, buf ^= 1) {
        const int k = t + threadIdx.x;
        tile[buf][threadIdx.x] = k < len ? src[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const int cnt = min(kBlock, len - t);
        const float4* t4 = tile[buf];
        for (int q = 0; q < cnt; ++q) {
            const float4 b = t4[q];
#pragma unroll
            for (int u = 0; u < I; ++u) {
                const float dx = b.x - me.x[u];
                const float dy = b.y - me.y[u];
                const float dz = b.z - me.z[u];
                float d2 = __builtin_fmaf(dx, dx, p.eps2);
                d2 = __builtin_fmaf(dy, dy, d2);
                d2 = __builtin_fmaf(dz, dz, d2);
                const float inv = rsqrt_f32<NEWTON>(d2);
                const double w = (double)((b.w * inv) * (inv * inv));
                acc.x[u] = __builtin_fma((double)dx, w, acc.x[u]);
                acc.y[u] = __builtin_fma((double)dy, w, acc.y[u]);
                acc.z[u] = __builtin_fma((double)dz, w, acc.z[u]);
            }
        }
    }
    const int mbase = blockIdx.x * (kBlock * I) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < I; ++u) {
        const int m = mbase + u * kBlock;
        if (m >= p.ni) continue;
        if (p.hi64 == nullptr) {
            float* f = p.force + 4 * (size_t)m;
            f[0] = (float)acc.x[u];
            f[1] = (float)acc.y[u];
            f[2] = (float)acc.z[u];
        } else {
            double* o = p.hi64 + (size_t)blockIdx.y * 3 * p.ni + m;
            o[0] = acc.x[u];
            o[p.ni] = acc.y[u];
            o[2 * (size_t)p.ni] = acc.z[u];
        }
    }
}

// The wave64 counterpart of the reference's warp-shuffle kernel
// (src/nbody_GPU_Shuffle.cu:44-89), where each lane loads one j body and the
// warp broadcasts every lane's body in turn: here each lane loads one body of
// a 64-body tile, every lane meets the body it holds with its I i bodies, and
// the tile then moves one lane along the wave (DPP wave_ror:1, one full-rate
// v_mov_b32_dpp per component), so after 64 steps every lane has met all 64.
// Per interaction that is the 12 VALU + 1 v_rsq_f32 chain plus 4/I moves;
// the row broadcast folded into the chain's operands (row_newbcast on three
// subtractions and a product, four DPP operands per interaction) measured
// 0.40 of the FP32 peak at N = 98304 against 0.46 for this. The lane holding
// body i loads tile body (i mod 64) first (shift = i0 mod 64 for the first
// body i0 of the range) and then the ones below it, wrapping around: a body's
// terms come in an order fixed by its own index, so an i-range evaluation
// (base, n) gives each body the bits of the whole-system one. Compensated
// plans fold the terms per 128 as the other variants do. Bodies past the
// last full tile go through the scalar-load sweep.
template <int I, bool NEWTON, bool EQ>
__device__ __forceinline__ void sweep_dpp(const float4* __restrict__ src, int len, int shift, float eps2,
                                          const Lanes<I>& me, Lanes<I>& acc)
{
    const int first = (int)(threadIdx.x + shift) & 63;
    int g = 0;
    for (; g + 64 <= len; g += 64) {
        const float4 b = src[g + first];
        float xj = b.x, yj = b.y, zj = b.z, mj = b.w;
        for (int k = 0; k < 64; ++k) {
#pragma unroll
            for (int u = 0; u < I; ++u)
                interact<NEWTON, false, EQ>(xj, yj, zj, mj, me.x[u], me.y[u], me.z[u], eps2, acc.x[u], acc.y[u],
                                            acc.z[u]);
            xj = wave_ror1(xj);
            yj = wave_ror1(yj);
            zj = wave_ror1(zj);
            if constexpr (!EQ) mj = wave_ror1(mj);
        }
    }
    // (through the constant address space: scalar loads, and no vector
    // registers for the group of j bodies in flight)
    if (g < len) sweep_const<I, NEWTON, EQ>(src + g, len - g, eps2, me, acc);
}

// Compensated plans keep the TwoSum residues of the DPP kernel in LDS, one
// column per thread (no thread reads another's, so no barrier), written once
// per kChunk terms: with 8 bodies per lane the residues in VGPRs took the
// kernel to 144 VGPRs, 3 waves per SIMD. Same arithmetic, same bits.
template <int I, bool COMP, bool NEWTON>
__global__ __launch_bounds__(kBlock) void k_force_dpp(ForceParams p)
{
    Lanes<I> me, hi;
    load_i<I>(p, me);
    zero<I>(hi);
    const int j0 = blockIdx.y * p.jslice;
    const int len = min(p.n - j0, p.jslice);
    const float4* __restrict__ src = p.pm + j0;
    const float eps2 = p.eps2;
    const int shift = p.i0 & 63;
    auto eq_sweep = [&](int c, int l, Lanes<I>& a) { sweep_dpp<I, NEWTON, true>(src + c, l, shift, eps2, me, a); };
    auto gen_sweep = [&](int c, int l, Lanes<I>& a) { sweep_dpp<I, NEWTON, false>(src + c, l, shift, eps2, me, a); };
    if constexpr (!COMP) {
        Lanes<I> lo;
        zero<I>(lo);
        slice_sums_mass<I, false>(src, len, eq_sweep, gen_sweep, hi, lo);
        store<I>(p, hi, nullptr);
    } else {
        __shared__ float res[3 * I][kBlock];
        const int t = threadIdx.x;
#pragma unroll
        for (int r = 0; r < 3 * I; ++r) res[r][t] = 0.f;
        auto sums = [&](const auto& sweep) {
            for (int c = 0; c < len; c += kChunk) {
                Lanes<I> part;
                zero<I>(part);
                sweep(c, min(kChunk, len - c), part);
#pragma unroll
                for (int u = 0; u < I; ++u) {
                    float e;
                    hi.x[u] = two_sum(hi.x[u], part.x[u], e);
                    res[u][t] += e;
                    hi.y[u] = two_sum(hi.y[u], part.y[u], e);
                    res[I + u][t] += e;
                    hi.z[u] = two_sum(hi.z[u], part.z[u], e);
                    res[2 * I + u][t] += e;
                }
            }
        };
        float m0;
        const bool eq = uniform_mass(src, len, m0);
        if (eq) sums(eq_sweep);
        else sums(gen_sweep);
        Lanes<I> lo;
#pragma unroll
        for (int u = 0; u < I; ++u) {
            lo.x[u] = res[u][t];
            lo.y[u] = res[I + u][t];
            lo.z[u] = res[2 * I + u][t];
        }
        if (eq) {
            scale<I>(hi, m0);
            scale<I>(lo, m0);
        }
        store<I>(p, hi, &lo);
    }
}

// ---- symmetric: each pair once, reaction kept on chip ----------------------
//
// The symmetric plan (see nbody_sym.h): workgroup t < nb(nb-1)/2 computes
// block pair (P, Q), P < Q; the remaining nb*A workgroups are the diagonal
// pieces, one-sided, last in the grid: their small size lets them fill the
// end of the launch evenly (one task per diagonal block, each pair once,
// halves their work but measured 2.5 % slower at N = 98304 for its coarser
// end; the batched kernel, whose queue mixes many systems, uses it).
// One launch covers the column band q0 <= Q < q1 of the block triangle: the
// pairs (P, Q), P < Q, then A pieces of each band block's diagonal.
struct SymParams {
    const float4* pm;
    int n;
    int q0;        // band: first column block
    int npairs;    // pairs in the band: Q(Q-1)/2 summed over its columns
    float eps2;
    SlotMap slots; // where the band's partials go
};

template <int A, int W, bool NEWTON>
__global__ __launch_bounds__(64 * W) void k_force_sym(SymParams p)
{
    constexpr int B = 64 * A * W;
    __shared__ float jacc[3][B];
    // Registers decide how many waves share a SIMD; LDS only how many
    // workgroups share a CU, which the dispatcher may spread unevenly over its
    // four SIMDs. With 8 bodies per lane on one wave the kernel needs 71
    // VGPRs (7 waves per SIMD; 6.5 by LDS) but runs fastest at exactly 6 per
    // SIMD (121 TFLOP/s credited at N = 98304 against 115), so it claims v79:
    // 80 VGPRs, 6 waves on every SIMD.
    if constexpr (A == 8 && W == 1) asm volatile("" ::: "v79");
    const int t = blockIdx.x;
    if (t < p.npairs) {
        int P, Q;
        upper_pair(p.q0 * (p.q0 - 1) / 2 + t, P, Q);
        sym_pair<A, W, NEWTON>(p.pm, p.n, p.eps2, p.slots, P, Q, jacc);
    } else {
        const int d = t - p.npairs;
        sym_diag_piece<B, NEWTON, false>(p.pm, p.n, p.eps2, p.slots, p.q0 + d / A, d % A, 64 * W);
    }
}

// The band's slots of each body summed in slot order with TwoSum. One lane
// per (body, component) (blockIdx.y = component); slots are fetched eight at
// a time so the loads overlap, then folded strictly in slot order. With one
// band the sum is rounded once into the force. With several, the running
// (sum, error) pair of each body carries from band to band in `state`
// ([2][3][n]; the first band starts it, a body of a block past the band adds
// nothing yet) and the last band rounds it into the force; each body's slots
// are then folded in band order, the same fixed order on every run. With
// force64 the last band writes the pair's value in FP64 instead (a share of
// the columns, whose partial force is summed with the other shares').
__global__ __launch_bounds__(kBlock) void k_reduce_band(SlotMap m, int n, int B, int q0, int q1, bool first,
                                                        bool last, float* __restrict__ state,
                                                        float* __restrict__ force, double* __restrict__ force64)
{
    constexpr int G = 8;
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= n) return;
    const int c = blockIdx.y;
    const int blk = b / B;
    if (!first && !last && blk >= q1) return;
    const float* src = nullptr;
    int slots = 0;
    size_t row = 0;
    if (blk < q0) {
        src = m.a + c * m.aplane + b;
        slots = q1 - q0;
        row = m.arow;
    } else if (blk < q1) {
        src = m.b + c * m.bplane + (b - (size_t)q0 * B);
        slots = q1;
        row = m.brow;
    }
    float s = 0.f, e = 0.f;
    if (!first) {
        s = state[(size_t)c * n + b];
        e = state[(size_t)(3 + c) * n + b];
    }
    int k = 0;
    for (; k + G <= slots; k += G) {
        float v[G];
#pragma unroll
        for (int g = 0; g < G; ++g) v[g] = src[(size_t)(k + g) * row];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            float err;
            s = two_sum(s, v[g], err);
            e += err;
        }
    }
    for (; k < slots; ++k) {
        float err;
        s = two_sum(s, src[(size_t)k * row], err);
        e += err;
    }
    if (last && force64) {
        force64[4 * (size_t)b + c] = (double)s + (double)e;
    } else if (last) {
        force[4 * (size_t)b + c] = s + e;
    } else {
        state[(size_t)c * n + b] = s;
        state[(size_t)(3 + c) * n + b] = e;
    }
}

// The symmetric plan with every term added in FP64 (the accurate plan): one
// wave per task, A bodies per lane, blocks of B = 64 A; slots are doubles.
struct SymParams64 {
    const float4* pm;
    int n;
    int q0;
    int npairs;
    float eps2;
    SlotMap64 slots;
};

template <int A, bool NEWTON>
__global__ __launch_bounds__(64) void k_force_sym64(SymParams64 p)
{
    const int t = blockIdx.x;
    if (t < p.npairs) {
        int P, Q;
        upper_pair(p.q0 * (p.q0 - 1) / 2 + t, P, Q);
        sym_pair64<A, NEWTON>(p.pm, p.n, p.eps2, p.slots, P, Q);
    } else {
        const int d = t - p.npairs;
        sym_diag_piece64<64 * A, NEWTON>(p.pm, p.n, p.eps2, p.slots, p.q0 + d / A, d % A);
    }
}

// k_reduce_band for FP64 slots: each body's slots of the band added in slot
// order in FP64, the running sum carried between bands in `state` ([3][n]
// doubles), rounded once into the force by the last band.
__global__ __launch_bounds__(kBlock) void k_reduce_band64(SlotMap64 m, int n, int B, int q0, int q1, bool first,
                                                          bool last, double* __restrict__ state,
                                                          float* __restrict__ force, double* __restrict__ force64)
{
    constexpr int G = 8;
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= n) return;
    const int c = blockIdx.y;
    const int blk = b / B;
    if (!first && !last && blk >= q1) return;
    const double* src = nullptr;
    int slots = 0;
    size_t row = 0;
    if (blk < q0) {
        src = m.a + c * m.aplane + b;
        slots = q1 - q0;
        row = m.arow;
    } else if (blk < q1) {
        src = m.b + c * m.bplane + (b - (size_t)q0 * B);
        slots = q1;
        row = m.brow;
    }
    double s = first ? 0.0 : state[(size_t)c * n + b];
    int k = 0;
    for (; k + G <= slots; k += G) {
        double v[G];
#pragma unroll
        for (int g = 0; g < G; ++g) v[g] = src[(size_t)(k + g) * row];
#pragma unroll
        for (int g = 0; g < G; ++g) s += v[g];
    }
    for (; k < slots; ++k) s += src[(size_t)k * row];
    if (last && force64) force64[4 * (size_t)b + c] = s;
    else if (last) force[4 * (size_t)b + c] = (float)s;
    else state[(size_t)c * n + b] = s;
}

// FP64 slice partials [slice][component][ni] added in slice order, rounded once.
__global__ __launch_bounds__(kBlock) void k_reduce64(const double* __restrict__ hi64, int nslices, int ni,
                                                     float* __restrict__ force)
{
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= ni) return;
    double x = 0.0, y = 0.0, z = 0.0;
    for (int s = 0; s < nslices; ++s) {
        const double* o = hi64 + (size_t)s * 3 * ni + m;
        x += o[0];
        y += o[ni];
        z += o[2 * (size_t)ni];
    }
    float* f = force + 4 * (size_t)m;
    f[0] = (float)x;
    f[1] = (float)y;
    f[2] = (float)z;
}

// Slice partials summed in slice order; compensated plans merge the TwoSum
// residues and round once at the end.
template <bool COMP>
__global__ __launch_bounds__(kBlock) void k_reduce(const float4* __restrict__ hi, const float4* __restrict__ lo,
                                                   int nslices, int ni, float* __restrict__ force)
{
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= ni) return;
    float4 a = hi[m];
    float4 c = COMP ? lo[m] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s = 1; s < nslices; ++s) {
        const size_t k = (size_t)s * ni + m;
        const float4 h = hi[k];
        if constexpr (COMP) {
            const float4 l = lo[k];
            float e;
            a.x = two_sum(a.x, h.x, e);
            c.x += e + l.x;
            a.y = two_sum(a.y, h.y, e);
            c.y += e + l.y;
            a.z = two_sum(a.z, h.z, e);
            c.z += e + l.z;
        } else {
            a.x += h.x;
            a.y += h.y;
            a.z += h.z;
        }
    }
    float* f = force + 4 * (size_t)m;
    f[0] = a.x + c.x;
    f[1] = a.y + c.y;
    f[2] = a.z + c.z;
}

// src/nbody.cu:158-182 integrateOne, in the operation order the reference's
// -ffast-math CPU build executes it: v = fma(invMass, f*dt, v); v *= damping;
// x = fma(v, dt, x). Mass (posMass.w) and invMass (velInvMass.w) are kept.
__global__ __launch_bounds__(kBlock) void k_integrate(float4* __restrict__ pm, float4* __restrict__ vim,
                                                      const float4* __restrict__ f, float dt, float damping, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float4 p = pm[i];
    float4 v = vim[i];
    const float4 g = f[i];
    v.x = __builtin_fmaf(v.w, g.x * dt, v.x);
    v.y = __builtin_fmaf(v.w, g.y * dt, v.y);
    v.z = __builtin_fmaf(v.w, g.z * dt, v.z);
    v.x *= damping;
    v.y *= damping;
    v.z *= damping;
    p.x = __builtin_fmaf(v.x, dt, p.x);
    p.y = __builtin_fmaf(v.y, dt, p.y);
    p.z = __builtin_fmaf(v.z, dt, p.z);
    pm[i] = p;
    vim[i] = v;
}

// FP64 direct sum, one i body per lane over one j slice (the j body again
// comes from scalar loads); IEEE division and square root.
__global__ __launch_bounds__(kBlock) void k_force_fp64(const float4* __restrict__ pm, int n, int i0, int ni,
                                                       double eps2, int jslice, double* __restrict__ partial)
{
    const int m = blockIdx.x * kBlock + threadIdx.x;
    const float4 q = pm[i0 + min(m, ni - 1)];
    const double xi = q.x, yi = q.y, zi = q.z;
    const int j0 = blockIdx.y * jslice;
    const int len = min(n - j0, jslice);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int j = 0; j < len; ++j) {
        const float4 b = pm[j0 + j];
        const double dx = (double)b.x - xi, dy = (double)b.y - yi, dz = (double)b.z - zi;
        const double d2 = dx * dx + dy * dy + dz * dz + eps2;
        const double inv = 1.0 / sqrt(d2);
        const double s = (double)b.w * (inv * inv * inv);
        ax += dx * s;
        ay += dy * s;
        az += dz * s;
    }
    if (m < ni) {
        double* o = partial + 3 * ((size_t)blockIdx.y * ni + m);
        o[0] = ax;
        o[1] = ay;
        o[2] = az;
    }
}

__device__ __forceinline__ void neumaier(double& s, double& c, double x)
{
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

__global__ __launch_bounds__(kBlock) void k_reduce_fp64(const double* __restrict__ partial, int nslices, int ni,
                                                        double* __restrict__ out)
{
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= ni) return;
    for (int c = 0; c < 3; ++c) {
        double s = 0.0, e = 0.0;
        for (int k = 0; k < nslices; ++k) neumaier(s, e, partial[3 * ((size_t)k * ni + m) + c]);
        out[4 * (size_t)m + c] = s + e;
    }
}

// Energy diagnostic, pass 1: block b reduces bodies [256b, 256b+256):
// kinetic 1/2 m v^2, momentum m v, and 1/2 sum_{j != i} -m_i m_j / sqrt(d^2 + eps2)
// (each pair seen from both ends, hence the half). Tree sums in LDS, fixed order.
__global__ __launch_bounds__(kBlock) void k_energy_fp64(const float4* __restrict__ pm, const float4* __restrict__ vim,
                                                        int n, double eps2, double* __restrict__ partial)
{
    __shared__ double red[5][kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const float4 q = pm[i];
        const float4 w = vim[i];
        const double m = q.w;
        v[0] = 0.5 * m * ((double)w.x * w.x + (double)w.y * w.y + (double)w.z * w.z);
        v[2] = m * w.x;
        v[3] = m * w.y;
        v[4] = m * w.z;
        double s = 0.0, e = 0.0;
        for (int j = 0; j < n; ++j) {
            const float4 b = pm[j];
            const double dx = (double)b.x - q.x, dy = (double)b.y - q.y, dz = (double)b.z - q.z;
            const double t = j == i ? 0.0 : (double)b.w / sqrt(dx * dx + dy * dy + dz * dz + eps2);
            neumaier(s, e, t);
        }
        v[1] = -0.5 * m * (s + e);
    }
    for (int c = 0; c < 5; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 5; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 5; ++c) partial[5 * (size_t)blockIdx.x + c] = red[c][0];
}

__global__ void k_energy_fp64_final(const double* __restrict__ partial, int blocks, double* __restrict__ acc)
{
    const int c = threadIdx.x;
    if (c >= 5) return;
    double s = 0.0, e = 0.0;
    for (int b = 0; b < blocks; ++b) neumaier(s, e, partial[5 * (size_t)b + c]);
    acc[c] = s + e;
}

// MAX_I: the most i bodies per lane a family is built for (8 only where the
// j body costs an instruction per step to reach the lanes: the DPP rotation).
#define NBODY_KERNEL_FAMILY(NAME, KERNEL, MAX_I)                                                 \
    struct NAME {                                                                                \
        static constexpr int kMaxI = MAX_I;                                                      \
        template <int I, bool C, bool N>                                                         \
        static void go(dim3 g, hipStream_t s, const ForceParams& p)                              \
        {                                                                                        \
            KERNEL<I, C, N><<<g, kBlock, 0, s>>>(p);                                             \
        }                                                                                        \
    };
NBODY_KERNEL_FAMILY(FamSgpr, k_force_sgpr, 4)
NBODY_KERNEL_FAMILY(FamLds, k_force_lds, 4)
NBODY_KERNEL_FAMILY(FamDpp, k_force_dpp, 8)
#undef NBODY_KERNEL_FAMILY

template <class Fam, int I>
void go_i(const ForcePlan& plan, dim3 g, hipStream_t s, const ForceParams& p)
{
    if (plan.comp) {
        if (plan.newton) Fam::template go<I, true, true>(g, s, p);
        else Fam::template go<I, true, false>(g, s, p);
    } else {
        if (plan.newton) Fam::template go<I, false, true>(g, s, p);
        else Fam::template go<I, false, false>(g, s, p);
    }
}

template <class Fam>
hipError_t go_family(const ForcePlan& plan, dim3 g, hipStream_t s, const ForceParams& p)
{
    switch (plan.ipl) {
    case 1: go_i<Fam, 1>(plan, g, s, p); break;
    case 2: go_i<Fam, 2>(plan, g, s, p); break;
    case 4: go_i<Fam, 4>(plan, g, s, p); break;
    case 8:
        if constexpr (Fam::kMaxI >= 8) {
            go_i<Fam, 8>(plan, g, s, p);
            break;
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Resident workgroups per CU of `kernel` with `dyn` bytes of dynamic LDS.
int blocks_per_cu(const void* kernel, int threads, size_t dyn)
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, dyn) != hipSuccess) return 0;
    return blocks;
}

// Smallest extra (dynamic) LDS per workgroup that keeps at most `waves`
// waves of the kernel on each SIMD (4 SIMDs per CU); 0 when the registers
// already keep it there. Cached per (kernel, waves).
size_t lds_for_occupancy(const void* kernel, int threads, int waves)
{
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, size_t> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(kernel, waves);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const int target = std::max(1, 4 * waves / (threads / 64));
    size_t dyn = 0;
    while (dyn < 160 * 1024 && blocks_per_cu(kernel, threads, dyn) > target) dyn += 256;
    cache[key] = dyn;
    return dyn;
}

template <int A, int W, bool NEWTON>
const void* sym_kernel()
{
    return reinterpret_cast<const void*>(&k_force_sym<A, W, NEWTON>);
}

template <int A, int W>
hipError_t go_sym(const ForcePlan& plan, const SymParams& sp, int band_blocks, hipStream_t s)
{
    const int tasks = sp.npairs + band_blocks * A;
    const void* k = plan.newton ? sym_kernel<A, W, true>() : sym_kernel<A, W, false>();
    const size_t dyn = plan.occ > 0 ? lds_for_occupancy(k, 64 * W, plan.occ) : 0;
    if (plan.newton) k_force_sym<A, W, true><<<tasks, 64 * W, dyn, s>>>(sp);
    else k_force_sym<A, W, false><<<tasks, 64 * W, dyn, s>>>(sp);
    return hipGetLastError();
}

hipError_t go_sym_shape(const ForcePlan& plan, const SymParams& sp, int band_blocks, hipStream_t s)
{
    const int w = plan.jslice / (64 * plan.ipl);
    if (plan.ipl == 2) return w == 1 ? go_sym<2, 1>(plan, sp, band_blocks, s) : go_sym<2, 4>(plan, sp, band_blocks, s);
    if (plan.ipl == 4)
        return w == 1   ? go_sym<4, 1>(plan, sp, band_blocks, s)
               : w == 2 ? go_sym<4, 2>(plan, sp, band_blocks, s)
                        : go_sym<4, 4>(plan, sp, band_blocks, s);
    return w == 1   ? go_sym<8, 1>(plan, sp, band_blocks, s)
           : w == 2 ? go_sym<8, 2>(plan, sp, band_blocks, s)
                    : go_sym<8, 4>(plan, sp, band_blocks, s);
}

// Floats of slot regions band [q0, q1) keeps: region A then region B.
size_t band_floats(int q0, int q1, int B) { return (size_t)3 * (q1 - q0) * B * ((size_t)q0 + q1); }

// Where band [q0, q1) keeps its slots, from `base`: region A, then region B.
template <class T>
SlotMapT<T> band_slots(T* base, int q0, int q1, int B)
{
    const int w = q1 - q0;
    SlotMapT<T> m;
    m.a = base;
    m.arow = q0 * B;
    m.aplane = (size_t)w * q0 * B;
    m.b = base + 3 * m.aplane;
    m.brow = w * B;
    m.bplane = (size_t)q1 * w * B;
    m.q0 = q0;
    return m;
}

template <int A, int W>
int natural_occupancy(bool newton)
{
    return blocks_per_cu(newton ? sym_kernel<A, W, true>() : sym_kernel<A, W, false>(), 64 * W, 0) * W / 4;
}

}  // namespace

bool symmetric_shape_ok(int ipl, int block)
{
    const int w = block / (64 * ipl);
    if (w * 64 * ipl != block) return false;
    return (ipl == 2 && (w == 1 || w == 4)) || (ipl == 4 && (w == 1 || w == 2 || w == 4)) ||
           (ipl == 8 && (w == 1 || w == 2 || w == 4));
}

int symmetric_natural_occupancy(const ForcePlan& plan)
{
    if (plan.variant != VAR_SYMMETRIC || !symmetric_shape_ok(plan.ipl, plan.jslice)) return 0;
    const int w = plan.jslice / (64 * plan.ipl);
    if (plan.ipl == 2) return w == 1 ? natural_occupancy<2, 1>(plan.newton) : natural_occupancy<2, 4>(plan.newton);
    if (plan.ipl == 4)
        return w == 1 ? natural_occupancy<4, 1>(plan.newton) : w == 2 ? natural_occupancy<4, 2>(plan.newton)
                                                                      : natural_occupancy<4, 4>(plan.newton);
    return w == 1 ? natural_occupancy<8, 1>(plan.newton) : w == 2 ? natural_occupancy<8, 2>(plan.newton)
                                                                  : natural_occupancy<8, 4>(plan.newton);
}

void symmetric_share_columns(int nb, int k, int G, int* c0, int* c1)
{
    // work of the columns below c: sum of (Q + 1) over Q < c = c (c + 1) / 2
    const double total = 0.5 * nb * (nb + 1.0);
    auto boundary = [&](int j) {
        if (j <= 0) return 0;
        if (j >= G) return nb;
        const double want = total * j / G;
        int c = (int)std::floor(std::sqrt(2.0 * want));
        while (c > 0 && 0.5 * c * (c + 1.0) > want) --c;
        while (c < nb && 0.5 * c * (c + 1.0) < want) ++c;
        return c;
    };
    *c0 = boundary(k);
    *c1 = boundary(k + 1);
}

std::vector<std::pair<int, int>> symmetric_bands(int nb, int B, size_t cap, int elem, int c0, int c1)
{
    std::vector<std::pair<int, int>> bands;
    if (c1 < 0) c1 = nb;
    const double per_col = 3.0 * elem * B;  // bytes per band column and block below q1, both regions
    for (int q0 = c0; q0 < c1;) {
        // the widest band [q0, q0 + w) whose regions, 3 elem B w (2 q0 + w) bytes, fit the cap
        const double w = -(double)q0 + std::sqrt((double)q0 * q0 + (double)cap / per_col);
        int width = std::max(1, std::min(c1 - q0, (int)w));
        while (width > 1 && band_floats(q0, q0 + width, B) * elem > cap) --width;
        bands.emplace_back(q0, q0 + width);
        q0 += width;
    }
    return bands;
}

size_t plan_scratch_bytes(const ForcePlan& plan, int ni)
{
    if (plan.variant == VAR_SYMMETRIC) {
        const int elem = plan.acc64 ? 8 : 4;
        const auto bands = symmetric_bands(plan.nslices, plan.jslice, plan.band_cap ? plan.band_cap : kSymBandCap,
                                           elem, plan.col0, plan.col1 > 0 ? plan.col1 : plan.nslices);
        size_t most = 0;
        for (const auto& b : bands) most = std::max(most, band_floats(b.first, b.second, plan.jslice));
        // the running sums carried between bands: (sum, error) float pairs or one double per component
        const size_t carry = bands.size() > 1 ? (size_t)6 * ni * sizeof(float) : 0;
        return most * elem + carry;
    }
    if (plan.nslices <= 1) return 0;
    if (plan.acc64) return (size_t)plan.nslices * ni * 3 * sizeof(double);
    return (size_t)plan.nslices * ni * sizeof(float4) * (plan.comp ? 2 : 1);
}

hipError_t launch_force(const ForcePlan& plan, const ForceParams& p, hipStream_t stream, hipEvent_t before,
                        hipEvent_t after)
{
    if (p.ni <= 0 || p.n <= 0) return hipSuccess;
    if (before) {
        hipError_t e = hipEventRecord(before, stream);
        if (e != hipSuccess) return e;
    }
    hipError_t e;
    if (plan.variant == VAR_AOS) {
        const dim3 g((p.ni + kBlock - 1) / kBlock, 1);
        ForceParams q = p;
        q.hi = q.lo = nullptr;
        if (plan.newton) k_force_aos<true><<<g, kBlock, 0, stream>>>(q);
        else k_force_aos<false><<<g, kBlock, 0, stream>>>(q);
        e = hipGetLastError();
        if (e == hipSuccess && after) e = hipEventRecord(after, stream);
        return e;
    }
    if (plan.variant == VAR_SYMMETRIC) {
        const bool ok_shape = plan.acc64 ? plan.jslice == 64 * plan.ipl && (plan.ipl == 4 || plan.ipl == 8)
                                         : symmetric_shape_ok(plan.ipl, plan.jslice);
        const int B = plan.jslice, nb = plan.nslices;
        const int c0 = plan.col0, c1 = plan.col1 > 0 ? plan.col1 : nb;
        // slots are needed unless the column range is empty (a share with no columns)
        const bool no_slots = plan.acc64 ? p.hi64 == nullptr : p.hi == nullptr;
        if (p.i0 != 0 || p.ni != p.n || !ok_shape || (no_slots && c1 > c0)) return hipErrorInvalidValue;
        const size_t cap = plan.band_cap ? plan.band_cap : kSymBandCap;
        auto bands = symmetric_bands(nb, B, cap, plan.acc64 ? 8 : 4, c0, c1);
        size_t most = 0;
        for (const auto& b : bands) most = std::max(most, band_floats(b.first, b.second, B));
        // a share with no columns still writes its (zero) partial forces
        if (bands.empty()) bands.emplace_back(c0, c0);
        const dim3 gr((p.n + kBlock - 1) / kBlock, 3);
        e = hipSuccess;
        for (size_t k = 0; k < bands.size(); ++k) {
            const int q0 = bands[k].first, q1 = bands[k].second, w = q1 - q0;
            const bool first = k == 0, last = k + 1 == bands.size();
            const int npairs = (int)(((long long)q1 * (q1 - 1) - (long long)q0 * (q0 - 1)) / 2);
            if (plan.acc64) {
                SymParams64 sp;
                sp.pm = p.pm;
                sp.n = p.n;
                sp.q0 = q0;
                sp.npairs = npairs;
                sp.eps2 = p.eps2;
                sp.slots = band_slots(p.hi64, q0, q1, B);
                const int tasks = npairs + w * plan.ipl;
                if (tasks > 0) {
                    if (plan.ipl == 4) {
                        if (plan.newton) k_force_sym64<4, true><<<tasks, 64, 0, stream>>>(sp);
                        else k_force_sym64<4, false><<<tasks, 64, 0, stream>>>(sp);
                    } else {
                        if (plan.newton) k_force_sym64<8, true><<<tasks, 64, 0, stream>>>(sp);
                        else k_force_sym64<8, false><<<tasks, 64, 0, stream>>>(sp);
                    }
                    e = hipGetLastError();
                }
                if (e == hipSuccess && after && last) e = hipEventRecord(after, stream);
                if (e != hipSuccess) return e;
                double* state = bands.size() > 1 ? p.hi64 + most : nullptr;
                k_reduce_band64<<<gr, kBlock, 0, stream>>>(sp.slots, p.n, B, q0, q1, first, last, state, p.force,
                                                           p.force64);
            } else {
                SymParams sp;
                sp.pm = p.pm;
                sp.n = p.n;
                sp.q0 = q0;
                sp.npairs = npairs;
                sp.eps2 = p.eps2;
                float* base = reinterpret_cast<float*>(p.hi);
                sp.slots = band_slots(base, q0, q1, B);
                if (npairs + w > 0) e = go_sym_shape(plan, sp, w, stream);
                if (e == hipSuccess && after && last) e = hipEventRecord(after, stream);
                if (e != hipSuccess) return e;
                float* state = bands.size() > 1 ? base + most : nullptr;
                k_reduce_band<<<gr, kBlock, 0, stream>>>(sp.slots, p.n, B, q0, q1, first, last, state, p.force,
                                                         p.force64);
            }
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const int per_block = kBlock * plan.ipl;
    const dim3 g((p.ni + per_block - 1) / per_block, p.nslices);
    if (plan.acc64) {
        if (plan.variant != VAR_LDS || (p.nslices > 1 && p.hi64 == nullptr)) return hipErrorInvalidValue;
        switch (plan.ipl) {
        case 1: plan.newton ? k_force_lds64<1, true><<<g, kBlock, 0, stream>>>(p)
                            : k_force_lds64<1, false><<<g, kBlock, 0, stream>>>(p); break;
        case 2: plan.newton ? k_force_lds64<2, true><<<g, kBlock, 0, stream>>>(p)
                            : k_force_lds64<2, false><<<g, kBlock, 0, stream>>>(p); break;
        case 4: plan.newton ? k_force_lds64<4, true><<<g, kBlock, 0, stream>>>(p)
                            : k_force_lds64<4, false><<<g, kBlock, 0, stream>>>(p); break;
        default: return hipErrorInvalidValue;
        }
        e = hipGetLastError();
        if (e == hipSuccess && after) e = hipEventRecord(after, stream);
        if (e != hipSuccess || p.hi64 == nullptr) return e;
        k_reduce64<<<(p.ni + kBlock - 1) / kBlock, kBlock, 0, stream>>>(p.hi64, p.nslices, p.ni, p.force);
        return hipGetLastError();
    }
    switch (plan.variant) {
    case VAR_SGPR: e = go_family<FamSgpr>(plan, g, stream, p); break;
    case VAR_LDS: e = go_family<FamLds>(plan, g, stream, p); break;
    case VAR_DPP: e = go_family<FamDpp>(plan, g, stream, p); break;
    default: return hipErrorNotSupported;
    }
    if (e == hipSuccess && after) e = hipEventRecord(after, stream);
    if (e != hipSuccess || p.hi == nullptr) return e;
    const dim3 gr((p.ni + kBlock - 1) / kBlock);
    if (plan.comp) k_reduce<true><<<gr, kBlock, 0, stream>>>(p.hi, p.lo, p.nslices, p.ni, p.force);
    else k_reduce<false><<<gr, kBlock, 0, stream>>>(p.hi, p.lo, p.nslices, p.ni, p.force);
    return hipGetLastError();
}

hipError_t launch_integrate(float4* posMass, float4* velInvMass, const float4* force, float dt, float damping, int n,
                            hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    k_integrate<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(posMass, velInvMass, force, dt, damping, n);
    return hipGetLastError();
}

hipError_t launch_force_fp64(double* out, const float4* pm, int n, int i0, int ni, double eps2, int jslice,
                             double* partial, hipStream_t stream)
{
    if (ni <= 0 || n <= 0) return hipSuccess;
    const int nslices = (n + jslice - 1) / jslice;
    k_force_fp64<<<dim3((ni + kBlock - 1) / kBlock, nslices), kBlock, 0, stream>>>(pm, n, i0, ni, eps2, jslice,
                                                                                  partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_fp64<<<(ni + kBlock - 1) / kBlock, kBlock, 0, stream>>>(partial, nslices, ni, out);
    return hipGetLastError();
}

int fp64_energy_blocks(int n) { return (n + kBlock - 1) / kBlock; }

hipError_t launch_energy_fp64(double* acc, const float4* pm, const float4* vim, int n, double eps2, double* partial,
                              hipStream_t stream)
{
    if (n <= 0) return hipErrorInvalidValue;
    const int blocks = fp64_energy_blocks(n);
    k_energy_fp64<<<blocks, kBlock, 0, stream>>>(pm, vim, n, eps2, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_energy_fp64_final<<<1, 64, 0, stream>>>(partial, blocks, acc);
    return hipGetLastError();
}

}  // namespace nbody
