// The symmetric decomposition's device code, shared by the single-system
// kernel (k_force_sym, nbody_kernels.hip) and the batched work-queue kernel
// (k_force_batch, nbody_batch.hip).
//
// Bodies are grouped in blocks of B = 64*A*W. A block pair (P, Q), P < Q, is
// one workgroup task: its W waves hold A bodies of block P per lane and stream
// block Q as 64-body tiles. Lane l loads one j body; after every step the j
// body's position, mass and reaction accumulator move one lane along the wave
// (DPP wave_ror:1), so in 64 steps each lane meets all 64 and the
// accumulators come back home. Per pair: 16 VALU + 1 v_rsq_f32 for both
// directions (the reaction on j carries m_i, where the reference's symmetric
// kernels use m_j: src/nbody_GPU_AOS_tiled.cu:133-141), or 14 VALU + 1
// v_rsq_f32 when both blocks hold bodies of one mass. Waves visit the tiles
// staggered (wave w takes tile (phase + w*A) mod T), so within a phase no two
// waves own a tile and the per-tile reaction sums enter LDS in a fixed order,
// with one barrier per phase: no atomics anywhere.
//
// The diagonal of block P (its bodies against each other) comes after every
// pair: in the single-system kernel as A one-sided pieces of 64*W bodies
// each (sym_diag_piece), in the batched kernel as one task that takes each
// pair of the block once (sym_diag_block, blocks of 512 on one wave).
//
// Every body then has exactly one partial per block slot: slot Q for the i
// side of (P, Q), slot P for its j side, slot P for the diagonal. The partials
// live in three float planes part[3][nb][npad] (npad = nb*B) and are added in
// slot order with TwoSum by the reduction kernels.
//
// A large system is evaluated in column bands of blocks Q in [q0, q1), one
// launch each, so that the partials of one band are all that is kept
// (SlotMap): a body of block P < q0 has one slot per band column (region A,
// [3][q1-q0][q0*B]); a body of a band block has one slot per block below q1
// (region B, [3][q1][(q1-q0)*B]): the j sides of the pairs (P, Q), its
// diagonal, and the i sides of its pairs with the band's later blocks. One
// band over all columns (q0 = 0) is exactly the layout above.
#pragma once

#include "nbody_device.h"

namespace nbody {

// Where the partials of one band go (see above), as elements of type T
// (float; double for the FP64-accumulation kernels). Plane c of a region
// starts c * plane elements after its base; a slot row holds `row` elements.
template <class T>
struct SlotMapT {
    T* a = nullptr;  // region A: bodies of blocks below q0, slot Q - q0
    T* b = nullptr;  // region B: bodies of blocks q0 .. q1-1, slot = the other block
    size_t aplane = 0, bplane = 0;
    int arow = 0, brow = 0;
    int q0 = 0;

    // One band over every column of a system: planes [3][nb][npad].
    __host__ __device__ static SlotMapT whole(T* part, int nb, int npad)
    {
        SlotMapT m;
        m.b = part;
        m.brow = npad;
        m.bplane = (size_t)nb * npad;
        return m;
    }
    // The i-side partials of block P's bodies against block Q (slot Q), B bodies per block.
    __device__ __forceinline__ T* i_side(int P, int Q, int B) const
    {
        return P < q0 ? a + (size_t)(Q - q0) * arow + (size_t)P * B : b + (size_t)Q * brow + (size_t)(P - q0) * B;
    }
    __device__ __forceinline__ size_t i_plane(int P) const { return P < q0 ? aplane : bplane; }
    // The j-side partials of block Q's bodies against block P (slot P); Q is a band block.
    __device__ __forceinline__ T* j_side(int P, int Q, int B) const
    {
        return b + (size_t)P * brow + (size_t)(Q - q0) * B;
    }
};
typedef SlotMapT<float> SlotMap;

// Every lane reads the lane below it, lane 0 reads lane 63 (DPP wave_ror:1).
// All source lanes are valid, so no "old" value has to be kept.
__device__ __forceinline__ float wave_ror1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13C, 0xf, 0xf, true));
}

// Task t < nb(nb-1)/2 of the strictly upper block triangle: t = Q(Q-1)/2 + P, P < Q.
__device__ __forceinline__ void upper_pair(int t, int& P, int& Q)
{
    int q = (int)((sqrtf(8.0f * (float)t + 1.0f) + 1.0f) * 0.5f);
    while (q > 1 && q * (q - 1) / 2 > t) --q;
    while ((q + 1) * q / 2 <= t) ++q;
    Q = q;
    P = t - q * (q - 1) / 2;
}

// One rotation step of a block pair with 8 i bodies per lane, in assembly so
// that its schedule is fixed: each pair's instructions run as one dependent
// chain (differences, d2, v_rsq_f32, the products, the reaction FMAs, then the
// FMAs into the i accumulators), the pairs one after another. Measured on
// gfx950 this order issues about 10% more VALU instructions per clock than
// any interleaving of independent pairs (the compiler's schedule interleaves
// them); the bank of each register hardly matters, so the compiler allocates
// them. One s_nop follows each v_rsq_f32 (the wait state gfx950 requires
// before a VALU reads a transcendental result), and the DPP rotations come
// four or more instructions after the last write of what they move. The
// arithmetic and its order are exactly those of the generic loop in sym_pair
// (same operations per accumulator, u in increasing order), so results are
// bit-identical to it. The i masses come in unnegated; the reaction weight
// -m_i s^3 takes its sign from the source modifier, so no negated copies of
// the masses need registers.
#define NBODY_SYM_PAIR(u)                                                                              \
    "v_sub_f32 %[dx], %[xj], %[x" #u "]\n"                                                             \
    "v_sub_f32 %[dy], %[yj], %[y" #u "]\n"                                                             \
    "v_sub_f32 %[dz], %[zj], %[z" #u "]\n"                                                             \
    "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                                           \
    "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                                  \
    "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                                  \
    "v_rsq_f32 %[s], %[s]\n"                                                                           \
    "s_nop 0\n"                                                                                        \
    "v_mul_f32 %[t], %[s], %[s]\n"                                                                     \
    "v_mul_f32 %[t], %[t], %[s]\n"                                                                     \
    "v_mul_f32 %[s], %[mj], %[t]\n"                                                                    \
    "v_mul_f32 %[t], -%[n" #u "], %[t]\n"                                                               \
    "v_fmac_f32 %[rx], %[dx], %[t]\n"                                                                  \
    "v_fmac_f32 %[ry], %[dy], %[t]\n"                                                                  \
    "v_fmac_f32 %[rz], %[dz], %[t]\n"                                                                  \
    "v_fmac_f32 %[ax" #u "], %[dx], %[s]\n"                                                            \
    "v_fmac_f32 %[ay" #u "], %[dy], %[s]\n"                                                            \
    "v_fmac_f32 %[az" #u "], %[dz], %[s]\n"

#define NBODY_ROR1(r) "v_mov_b32_dpp %[" #r "], %[" #r "] wave_ror:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"

// The first pair of a step whose reactions are summed apart (the batched
// kernel's steps, below): it starts the step's reaction partial with v_mul
// instead of adding to the running reaction; NBODY_SYM_BLOCK_ADD then adds
// the partial of the step's 8 pairs to the running reaction (Rx, Ry, Rz).
#define NBODY_SYM_PAIR_EQ_FIRST(u)                                                                     \
    "v_sub_f32 %[dx], %[xj], %[x" #u "]\n"                                                             \
    "v_sub_f32 %[dy], %[yj], %[y" #u "]\n"                                                             \
    "v_sub_f32 %[dz], %[zj], %[z" #u "]\n"                                                             \
    "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                                           \
    "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                                  \
    "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                                  \
    "v_rsq_f32 %[s], %[s]\n"                                                                           \
    "s_nop 0\n"                                                                                        \
    "v_mul_f32 %[t], %[s], %[s]\n"                                                                     \
    "v_mul_f32 %[t], %[t], %[s]\n"                                                                     \
    "v_mul_f32 %[rx], -%[dx], %[t]\n"                                                                  \
    "v_mul_f32 %[ry], -%[dy], %[t]\n"                                                                  \
    "v_mul_f32 %[rz], -%[dz], %[t]\n"                                                                  \
    "v_fmac_f32 %[ax" #u "], %[dx], %[t]\n"                                                            \
    "v_fmac_f32 %[ay" #u "], %[dy], %[t]\n"                                                            \
    "v_fmac_f32 %[az" #u "], %[dz], %[t]\n"

#define NBODY_SYM_PAIR_FIRST(u)                                                                        \
    "v_sub_f32 %[dx], %[xj], %[x" #u "]\n"                                                             \
    "v_sub_f32 %[dy], %[yj], %[y" #u "]\n"                                                             \
    "v_sub_f32 %[dz], %[zj], %[z" #u "]\n"                                                             \
    "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                                           \
    "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                                  \
    "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                                  \
    "v_rsq_f32 %[s], %[s]\n"                                                                           \
    "s_nop 0\n"                                                                                        \
    "v_mul_f32 %[t], %[s], %[s]\n"                                                                     \
    "v_mul_f32 %[t], %[t], %[s]\n"                                                                     \
    "v_mul_f32 %[s], %[mj], %[t]\n"                                                                    \
    "v_mul_f32 %[t], -%[n" #u "], %[t]\n"                                                               \
    "v_mul_f32 %[rx], %[dx], %[t]\n"                                                                   \
    "v_mul_f32 %[ry], %[dy], %[t]\n"                                                                   \
    "v_mul_f32 %[rz], %[dz], %[t]\n"                                                                   \
    "v_fmac_f32 %[ax" #u "], %[dx], %[s]\n"                                                            \
    "v_fmac_f32 %[ay" #u "], %[dy], %[s]\n"                                                            \
    "v_fmac_f32 %[az" #u "], %[dz], %[s]\n"

#define NBODY_SYM_BLOCK_ADD                                                                            \
    "v_add_f32 %[Rx], %[Rx], %[rx]\n"                                                                  \
    "v_add_f32 %[Ry], %[Ry], %[ry]\n"                                                                  \
    "v_add_f32 %[Rz], %[Rz], %[rz]\n"

// The chain's five temporaries (d, then s, then t = s^3) are either left to
// the compiler or, with PIN, held in v21, v25, v29, v33 and v34. On gfx950
// the issue rate of one instruction stream depends on which registers hold
// them relative to the rest: the loop bodies run verbatim, ALU only, issue
// 1.18 to 1.29 VALU instructions per clock per CU. The single-system kernel's
// own allocation measures at the top of that range. In the batched kernel
// the compiler picks a slow one, and pinning raises that kernel's rate by
// about 7 %.
// The batched kernel's steps (PIN) also sum each step's 8 reaction terms
// apart and add that partial to the running reaction (3 more VALU per 8
// pairs, 1.2 % of its rate): the running j-side sums then take 64 roundings
// per tile instead of 512, which keeps every system of the 512-system
// ensemble within 1e-5 of the FP64 sum on its evolved states (1.8e-5
// without). The single-system kernel keeps one running sum: it holds the
// bound on its configured state without the extra adds.
template <bool PIN>
__device__ __forceinline__ void sym_step8(float& xj, float& yj, float& zj, float& mj, float& rx, float& ry, float& rz,
                                          Lanes<8>& acc, const Lanes<8>& me, const float (&mi)[8], float eps2)
{
    float dx, dy, dz, s, t;
    if constexpr (PIN) {
        float qx, qy, qz;
        asm volatile(NBODY_SYM_PAIR_FIRST(0) NBODY_SYM_PAIR(1) NBODY_SYM_PAIR(2) NBODY_SYM_PAIR(3)
                     NBODY_SYM_PAIR(4) NBODY_SYM_PAIR(5) NBODY_SYM_PAIR(6) NBODY_SYM_PAIR(7)
                     NBODY_SYM_BLOCK_ADD
                     NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj) NBODY_ROR1(mj)
                     NBODY_ROR1(Rx) NBODY_ROR1(Ry) NBODY_ROR1(Rz)
                     : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [mj] "+v"(mj), [Rx] "+v"(rx), [Ry] "+v"(ry), [Rz] "+v"(rz),
                       [rx] "=&v"(qx), [ry] "=&v"(qy), [rz] "=&v"(qz),
                       [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ax2] "+v"(acc.x[2]),
                       [ax3] "+v"(acc.x[3]), [ax4] "+v"(acc.x[4]), [ax5] "+v"(acc.x[5]), [ax6] "+v"(acc.x[6]),
                       [ax7] "+v"(acc.x[7]), [ay0] "+v"(acc.y[0]), [ay1] "+v"(acc.y[1]), [ay2] "+v"(acc.y[2]),
                       [ay3] "+v"(acc.y[3]), [ay4] "+v"(acc.y[4]), [ay5] "+v"(acc.y[5]), [ay6] "+v"(acc.y[6]),
                       [ay7] "+v"(acc.y[7]), [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [az2] "+v"(acc.z[2]),
                       [az3] "+v"(acc.z[3]), [az4] "+v"(acc.z[4]), [az5] "+v"(acc.z[5]), [az6] "+v"(acc.z[6]),
                       [az7] "+v"(acc.z[7]), [dx] "=&{v21}"(dx), [dy] "=&{v25}"(dy), [dz] "=&{v29}"(dz),
                       [s] "=&{v33}"(s), [t] "=&{v34}"(t)
                     : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [x2] "v"(me.x[2]), [x3] "v"(me.x[3]),
                       [x4] "v"(me.x[4]), [x5] "v"(me.x[5]), [x6] "v"(me.x[6]), [x7] "v"(me.x[7]),
                       [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), [y2] "v"(me.y[2]), [y3] "v"(me.y[3]),
                       [y4] "v"(me.y[4]), [y5] "v"(me.y[5]), [y6] "v"(me.y[6]), [y7] "v"(me.y[7]),
                       [z0] "v"(me.z[0]), [z1] "v"(me.z[1]), [z2] "v"(me.z[2]), [z3] "v"(me.z[3]),
                       [z4] "v"(me.z[4]), [z5] "v"(me.z[5]), [z6] "v"(me.z[6]), [z7] "v"(me.z[7]),
                       [n0] "v"(mi[0]), [n1] "v"(mi[1]), [n2] "v"(mi[2]), [n3] "v"(mi[3]),
                       [n4] "v"(mi[4]), [n5] "v"(mi[5]), [n6] "v"(mi[6]), [n7] "v"(mi[7]), [eps] "s"(eps2));
    } else {
        asm volatile(NBODY_SYM_PAIR(0) NBODY_SYM_PAIR(1) NBODY_SYM_PAIR(2) NBODY_SYM_PAIR(3)
                     NBODY_SYM_PAIR(4) NBODY_SYM_PAIR(5) NBODY_SYM_PAIR(6) NBODY_SYM_PAIR(7)
                     NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj) NBODY_ROR1(mj)
                     NBODY_ROR1(rx) NBODY_ROR1(ry) NBODY_ROR1(rz)
                     : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [mj] "+v"(mj), [rx] "+v"(rx), [ry] "+v"(ry),
                       [rz] "+v"(rz), [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ax2] "+v"(acc.x[2]),
                       [ax3] "+v"(acc.x[3]), [ax4] "+v"(acc.x[4]), [ax5] "+v"(acc.x[5]), [ax6] "+v"(acc.x[6]),
                       [ax7] "+v"(acc.x[7]), [ay0] "+v"(acc.y[0]), [ay1] "+v"(acc.y[1]), [ay2] "+v"(acc.y[2]),
                       [ay3] "+v"(acc.y[3]), [ay4] "+v"(acc.y[4]), [ay5] "+v"(acc.y[5]), [ay6] "+v"(acc.y[6]),
                       [ay7] "+v"(acc.y[7]), [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [az2] "+v"(acc.z[2]),
                       [az3] "+v"(acc.z[3]), [az4] "+v"(acc.z[4]), [az5] "+v"(acc.z[5]), [az6] "+v"(acc.z[6]),
                       [az7] "+v"(acc.z[7]), [dx] "=&v"(dx), [dy] "=&v"(dy), [dz] "=&v"(dz), [s] "=&v"(s),
                       [t] "=&v"(t)
                     : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [x2] "v"(me.x[2]), [x3] "v"(me.x[3]),
                       [x4] "v"(me.x[4]), [x5] "v"(me.x[5]), [x6] "v"(me.x[6]), [x7] "v"(me.x[7]),
                       [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), [y2] "v"(me.y[2]), [y3] "v"(me.y[3]),
                       [y4] "v"(me.y[4]), [y5] "v"(me.y[5]), [y6] "v"(me.y[6]), [y7] "v"(me.y[7]),
                       [z0] "v"(me.z[0]), [z1] "v"(me.z[1]), [z2] "v"(me.z[2]), [z3] "v"(me.z[3]),
                       [z4] "v"(me.z[4]), [z5] "v"(me.z[5]), [z6] "v"(me.z[6]), [z7] "v"(me.z[7]),
                       [n0] "v"(mi[0]), [n1] "v"(mi[1]), [n2] "v"(mi[2]), [n3] "v"(mi[3]),
                       [n4] "v"(mi[4]), [n5] "v"(mi[5]), [n6] "v"(mi[6]), [n7] "v"(mi[7]), [eps] "s"(eps2));
    }
}

// The same step for a block pair whose bodies all have one mass m0: the pair
// weights are then +inv^3 (i side) and -inv^3 (j side) times m0, so the
// chain drops the two mass products (14 VALU + 1 v_rsq_f32 per pair) and the
// mass no longer rotates; sym_pair multiplies the partials by m0 when it
// stores them. The reaction FMA negates d through the source modifier.
// Bit-identical to the generic loop with EQ.
#define NBODY_SYM_PAIR_EQ(u)                                                                           \
    "v_sub_f32 %[dx], %[xj], %[x" #u "]\n"                                                             \
    "v_sub_f32 %[dy], %[yj], %[y" #u "]\n"                                                             \
    "v_sub_f32 %[dz], %[zj], %[z" #u "]\n"                                                             \
    "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                                           \
    "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                                  \
    "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                                  \
    "v_rsq_f32 %[s], %[s]\n"                                                                           \
    "s_nop 0\n"                                                                                        \
    "v_mul_f32 %[t], %[s], %[s]\n"                                                                     \
    "v_mul_f32 %[t], %[t], %[s]\n"                                                                     \
    "v_fma_f32 %[rx], -%[dx], %[t], %[rx]\n"                                                           \
    "v_fma_f32 %[ry], -%[dy], %[t], %[ry]\n"                                                           \
    "v_fma_f32 %[rz], -%[dz], %[t], %[rz]\n"                                                           \
    "v_fmac_f32 %[ax" #u "], %[dx], %[t]\n"                                                            \
    "v_fmac_f32 %[ay" #u "], %[dy], %[t]\n"                                                            \
    "v_fmac_f32 %[az" #u "], %[dz], %[t]\n"

// The pair chains with the weight s^3 cleared on the lanes whose bit in
// %[mask] is 0 (v_cndmask right after the cube): those lanes add exact zeros.
// Used for the one step of a diagonal tile against itself that would
// otherwise take each pair twice.
#define NBODY_SYM_PAIR_M(u)                                                                            \
    "v_sub_f32 %[dx], %[xj], %[x" #u "]\n"                                                             \
    "v_sub_f32 %[dy], %[yj], %[y" #u "]\n"                                                             \
    "v_sub_f32 %[dz], %[zj], %[z" #u "]\n"                                                             \
    "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                                           \
    "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                                  \
    "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                                  \
    "v_rsq_f32 %[s], %[s]\n"                                                                           \
    "s_nop 0\n"                                                                                        \
    "v_mul_f32 %[t], %[s], %[s]\n"                                                                     \
    "v_mul_f32 %[t], %[t], %[s]\n"                                                                     \
    "v_cndmask_b32 %[t], 0, %[t], %[mask]\n"                                                           \
    "v_mul_f32 %[s], %[mj], %[t]\n"                                                                    \
    "v_mul_f32 %[t], -%[n" #u "], %[t]\n"                                                               \
    "v_fmac_f32 %[rx], %[dx], %[t]\n"                                                                  \
    "v_fmac_f32 %[ry], %[dy], %[t]\n"                                                                  \
    "v_fmac_f32 %[rz], %[dz], %[t]\n"                                                                  \
    "v_fmac_f32 %[ax" #u "], %[dx], %[s]\n"                                                            \
    "v_fmac_f32 %[ay" #u "], %[dy], %[s]\n"                                                            \
    "v_fmac_f32 %[az" #u "], %[dz], %[s]\n"

#define NBODY_SYM_PAIR_EQ_M(u)                                                                         \
    "v_sub_f32 %[dx], %[xj], %[x" #u "]\n"                                                             \
    "v_sub_f32 %[dy], %[yj], %[y" #u "]\n"                                                             \
    "v_sub_f32 %[dz], %[zj], %[z" #u "]\n"                                                             \
    "v_fma_f32 %[s], %[dx], %[dx], %[eps]\n"                                                           \
    "v_fmac_f32 %[s], %[dy], %[dy]\n"                                                                  \
    "v_fmac_f32 %[s], %[dz], %[dz]\n"                                                                  \
    "v_rsq_f32 %[s], %[s]\n"                                                                           \
    "s_nop 0\n"                                                                                        \
    "v_mul_f32 %[t], %[s], %[s]\n"                                                                     \
    "v_mul_f32 %[t], %[t], %[s]\n"                                                                     \
    "v_cndmask_b32 %[t], 0, %[t], %[mask]\n"                                                           \
    "v_fma_f32 %[rx], -%[dx], %[t], %[rx]\n"                                                           \
    "v_fma_f32 %[ry], -%[dy], %[t], %[ry]\n"                                                           \
    "v_fma_f32 %[rz], -%[dz], %[t], %[rz]\n"                                                           \
    "v_fmac_f32 %[ax" #u "], %[dx], %[t]\n"                                                            \
    "v_fmac_f32 %[ay" #u "], %[dy], %[t]\n"                                                            \
    "v_fmac_f32 %[az" #u "], %[dz], %[t]\n"

template <bool PIN>
__device__ __forceinline__ void sym_step8_eq(float& xj, float& yj, float& zj, float& rx, float& ry, float& rz,
                                             Lanes<8>& acc, const Lanes<8>& me, float eps2)
{
    float dx, dy, dz, s, t;
    if constexpr (PIN) {
        float qx, qy, qz;
        asm volatile(NBODY_SYM_PAIR_EQ_FIRST(0) NBODY_SYM_PAIR_EQ(1) NBODY_SYM_PAIR_EQ(2) NBODY_SYM_PAIR_EQ(3)
                     NBODY_SYM_PAIR_EQ(4) NBODY_SYM_PAIR_EQ(5) NBODY_SYM_PAIR_EQ(6) NBODY_SYM_PAIR_EQ(7)
                     NBODY_SYM_BLOCK_ADD
                     NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj)
                     NBODY_ROR1(Rx) NBODY_ROR1(Ry) NBODY_ROR1(Rz)
                     : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [Rx] "+v"(rx), [Ry] "+v"(ry), [Rz] "+v"(rz),
                       [rx] "=&v"(qx), [ry] "=&v"(qy), [rz] "=&v"(qz),
                       [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ax2] "+v"(acc.x[2]),
                       [ax3] "+v"(acc.x[3]), [ax4] "+v"(acc.x[4]), [ax5] "+v"(acc.x[5]), [ax6] "+v"(acc.x[6]),
                       [ax7] "+v"(acc.x[7]), [ay0] "+v"(acc.y[0]), [ay1] "+v"(acc.y[1]), [ay2] "+v"(acc.y[2]),
                       [ay3] "+v"(acc.y[3]), [ay4] "+v"(acc.y[4]), [ay5] "+v"(acc.y[5]), [ay6] "+v"(acc.y[6]),
                       [ay7] "+v"(acc.y[7]), [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [az2] "+v"(acc.z[2]),
                       [az3] "+v"(acc.z[3]), [az4] "+v"(acc.z[4]), [az5] "+v"(acc.z[5]), [az6] "+v"(acc.z[6]),
                       [az7] "+v"(acc.z[7]), [dx] "=&{v21}"(dx), [dy] "=&{v25}"(dy), [dz] "=&{v29}"(dz),
                       [s] "=&{v33}"(s), [t] "=&{v34}"(t)
                     : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [x2] "v"(me.x[2]), [x3] "v"(me.x[3]),
                       [x4] "v"(me.x[4]), [x5] "v"(me.x[5]), [x6] "v"(me.x[6]), [x7] "v"(me.x[7]),
                       [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), [y2] "v"(me.y[2]), [y3] "v"(me.y[3]),
                       [y4] "v"(me.y[4]), [y5] "v"(me.y[5]), [y6] "v"(me.y[6]), [y7] "v"(me.y[7]),
                       [z0] "v"(me.z[0]), [z1] "v"(me.z[1]), [z2] "v"(me.z[2]), [z3] "v"(me.z[3]),
                       [z4] "v"(me.z[4]), [z5] "v"(me.z[5]), [z6] "v"(me.z[6]), [z7] "v"(me.z[7]), [eps] "s"(eps2));
    } else {
        asm volatile(NBODY_SYM_PAIR_EQ(0) NBODY_SYM_PAIR_EQ(1) NBODY_SYM_PAIR_EQ(2) NBODY_SYM_PAIR_EQ(3)
                     NBODY_SYM_PAIR_EQ(4) NBODY_SYM_PAIR_EQ(5) NBODY_SYM_PAIR_EQ(6) NBODY_SYM_PAIR_EQ(7)
                     NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj)
                     NBODY_ROR1(rx) NBODY_ROR1(ry) NBODY_ROR1(rz)
                     : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [rx] "+v"(rx), [ry] "+v"(ry),
                       [rz] "+v"(rz), [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ax2] "+v"(acc.x[2]),
                       [ax3] "+v"(acc.x[3]), [ax4] "+v"(acc.x[4]), [ax5] "+v"(acc.x[5]), [ax6] "+v"(acc.x[6]),
                       [ax7] "+v"(acc.x[7]), [ay0] "+v"(acc.y[0]), [ay1] "+v"(acc.y[1]), [ay2] "+v"(acc.y[2]),
                       [ay3] "+v"(acc.y[3]), [ay4] "+v"(acc.y[4]), [ay5] "+v"(acc.y[5]), [ay6] "+v"(acc.y[6]),
                       [ay7] "+v"(acc.y[7]), [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [az2] "+v"(acc.z[2]),
                       [az3] "+v"(acc.z[3]), [az4] "+v"(acc.z[4]), [az5] "+v"(acc.z[5]), [az6] "+v"(acc.z[6]),
                       [az7] "+v"(acc.z[7]), [dx] "=&v"(dx), [dy] "=&v"(dy), [dz] "=&v"(dz), [s] "=&v"(s),
                       [t] "=&v"(t)
                     : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [x2] "v"(me.x[2]), [x3] "v"(me.x[3]),
                       [x4] "v"(me.x[4]), [x5] "v"(me.x[5]), [x6] "v"(me.x[6]), [x7] "v"(me.x[7]),
                       [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), [y2] "v"(me.y[2]), [y3] "v"(me.y[3]),
                       [y4] "v"(me.y[4]), [y5] "v"(me.y[5]), [y6] "v"(me.y[6]), [y7] "v"(me.y[7]),
                       [z0] "v"(me.z[0]), [z1] "v"(me.z[1]), [z2] "v"(me.z[2]), [z3] "v"(me.z[3]),
                       [z4] "v"(me.z[4]), [z5] "v"(me.z[5]), [z6] "v"(me.z[6]), [z7] "v"(me.z[7]), [eps] "s"(eps2));
    }
}

// The same step for 2 bodies per lane (the smallest block shapes). With 4
// bodies per lane the compiler's schedule of the generic loop measured
// faster than this chain order (blocks of 256 at N = 24576..65536: 0.54-0.71
// of the FP32 peak against 0.53-0.68; the same at N = 16384), so that shape
// runs the generic loop; with 2 and 8 the chain order is faster.
__device__ __forceinline__ void sym_step2(float& xj, float& yj, float& zj, float& mj, float& rx, float& ry, float& rz,
                                          Lanes<2>& acc, const Lanes<2>& me, const float (&mi)[2], float eps2)
{
    float dx, dy, dz, s, t;
    asm volatile(NBODY_SYM_PAIR(0) NBODY_SYM_PAIR(1)
                 NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj) NBODY_ROR1(mj)
                 NBODY_ROR1(rx) NBODY_ROR1(ry) NBODY_ROR1(rz)
                 : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [mj] "+v"(mj), [rx] "+v"(rx), [ry] "+v"(ry),
                   [rz] "+v"(rz), [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ay0] "+v"(acc.y[0]),
                   [ay1] "+v"(acc.y[1]), [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [dx] "=&v"(dx),
                   [dy] "=&v"(dy), [dz] "=&v"(dz), [s] "=&v"(s), [t] "=&v"(t)
                 : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), [z0] "v"(me.z[0]),
                   [z1] "v"(me.z[1]), [n0] "v"(mi[0]), [n1] "v"(mi[1]), [eps] "s"(eps2));
}

__device__ __forceinline__ void sym_step2_eq(float& xj, float& yj, float& zj, float& rx, float& ry, float& rz,
                                          Lanes<2>& acc, const Lanes<2>& me, float eps2)
{
    float dx, dy, dz, s, t;
    asm volatile(NBODY_SYM_PAIR_EQ(0) NBODY_SYM_PAIR_EQ(1)
                 NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj)
                 NBODY_ROR1(rx) NBODY_ROR1(ry) NBODY_ROR1(rz)
                 : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [rx] "+v"(rx), [ry] "+v"(ry), [rz] "+v"(rz),
                   [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ay0] "+v"(acc.y[0]), [ay1] "+v"(acc.y[1]),
                   [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [dx] "=&v"(dx), [dy] "=&v"(dy), [dz] "=&v"(dz),
                   [s] "=&v"(s), [t] "=&v"(t)
                 : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), [z0] "v"(me.z[0]),
                   [z1] "v"(me.z[1]), [eps] "s"(eps2));
}

// One rotation step of diagonal tile T (see sym_diag_phases) in assembly, in
// the chain order of sym_step8: the pairs with the i sub-blocks u < T, then
// (MODE 1) the tile's own sub-block, or (MODE 2) the same with the weights of
// the lanes outside `mask` cleared; then the rotation. Every operand of
// sym_step8 is bound (they are live through the task anyway), so the
// registers stay those of the block-pair step; the temporaries are pinned as
// in sym_step8<true> (the batched kernel is the only user). With EQ the mass
// neither enters nor rotates.
#define NBODY_DIAG_OPS \
    : [xj] "+v"(xj), [yj] "+v"(yj), [zj] "+v"(zj), [mj] "+v"(mj), [rx] "+v"(rx), [ry] "+v"(ry), [rz] "+v"(rz), \
      [ax0] "+v"(acc.x[0]), [ax1] "+v"(acc.x[1]), [ax2] "+v"(acc.x[2]), [ax3] "+v"(acc.x[3]), \
      [ax4] "+v"(acc.x[4]), [ax5] "+v"(acc.x[5]), [ax6] "+v"(acc.x[6]), [ax7] "+v"(acc.x[7]), \
      [ay0] "+v"(acc.y[0]), [ay1] "+v"(acc.y[1]), [ay2] "+v"(acc.y[2]), [ay3] "+v"(acc.y[3]), \
      [ay4] "+v"(acc.y[4]), [ay5] "+v"(acc.y[5]), [ay6] "+v"(acc.y[6]), [ay7] "+v"(acc.y[7]), \
      [az0] "+v"(acc.z[0]), [az1] "+v"(acc.z[1]), [az2] "+v"(acc.z[2]), [az3] "+v"(acc.z[3]), \
      [az4] "+v"(acc.z[4]), [az5] "+v"(acc.z[5]), [az6] "+v"(acc.z[6]), [az7] "+v"(acc.z[7]), [dx] "=&{v21}"(dx), \
      [dy] "=&{v25}"(dy), [dz] "=&{v29}"(dz), [s] "=&{v33}"(s), [t] "=&{v34}"(t) \
    : [x0] "v"(me.x[0]), [x1] "v"(me.x[1]), [x2] "v"(me.x[2]), [x3] "v"(me.x[3]), [x4] "v"(me.x[4]), \
      [x5] "v"(me.x[5]), [x6] "v"(me.x[6]), [x7] "v"(me.x[7]), [y0] "v"(me.y[0]), [y1] "v"(me.y[1]), \
      [y2] "v"(me.y[2]), [y3] "v"(me.y[3]), [y4] "v"(me.y[4]), [y5] "v"(me.y[5]), [y6] "v"(me.y[6]), \
      [y7] "v"(me.y[7]), [z0] "v"(me.z[0]), [z1] "v"(me.z[1]), [z2] "v"(me.z[2]), [z3] "v"(me.z[3]), \
      [z4] "v"(me.z[4]), [z5] "v"(me.z[5]), [z6] "v"(me.z[6]), [z7] "v"(me.z[7]), [n0] "v"(mi[0]), \
      [n1] "v"(mi[1]), [n2] "v"(mi[2]), [n3] "v"(mi[3]), [n4] "v"(mi[4]), [n5] "v"(mi[5]), [n6] "v"(mi[6]), \
      [n7] "v"(mi[7]), [eps] "s"(eps2), [mask] "s"(mask)

// NBODY_LT<k>(P) expands to P(0) ... P(k-1): the pairs with the sub-blocks below tile k.
#define NBODY_LT0(P)
#define NBODY_LT1(P) NBODY_LT0(P) P(0)
#define NBODY_LT2(P) NBODY_LT1(P) P(1)
#define NBODY_LT3(P) NBODY_LT2(P) P(2)
#define NBODY_LT4(P) NBODY_LT3(P) P(3)
#define NBODY_LT5(P) NBODY_LT4(P) P(4)
#define NBODY_LT6(P) NBODY_LT5(P) P(5)
#define NBODY_LT7(P) NBODY_LT6(P) P(6)
#define NBODY_ROR_GEN \
    NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj) NBODY_ROR1(mj) NBODY_ROR1(rx) NBODY_ROR1(ry) NBODY_ROR1(rz)
#define NBODY_ROR_EQ NBODY_ROR1(xj) NBODY_ROR1(yj) NBODY_ROR1(zj) NBODY_ROR1(rx) NBODY_ROR1(ry) NBODY_ROR1(rz)

template <int T, int MODE, bool EQ>
__device__ __forceinline__ void sym_diag_step8(float& xj, float& yj, float& zj, float& mj, float& rx, float& ry,
                                               float& rz, Lanes<8>& acc, const Lanes<8>& me,
                                               const float (&mi)[8], float eps2, unsigned long long mask)
{
    float dx, dy, dz, s, t;
    // The rotation-only steps (tile 0 outside its own sub-block) begin with a
    // DPP read; its source may have been written by a VALU op just before
    // the block (the j load's select, the zeroed reactions), and gfx9 needs
    // two wait states between the two, which the compiler does not see
    // inside assembly: s_nop 1 provides them.
    if constexpr (!EQ && T == 0 && MODE == 0)
        asm volatile("s_nop 1\n" NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 0 && MODE == 1)
        asm volatile(NBODY_LT0(NBODY_SYM_PAIR) NBODY_SYM_PAIR(0) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 0 && MODE == 2)
        asm volatile(NBODY_LT0(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(0) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 1 && MODE == 0)
        asm volatile(NBODY_LT1(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 1 && MODE == 1)
        asm volatile(NBODY_LT1(NBODY_SYM_PAIR) NBODY_SYM_PAIR(1) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 1 && MODE == 2)
        asm volatile(NBODY_LT1(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(1) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 2 && MODE == 0)
        asm volatile(NBODY_LT2(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 2 && MODE == 1)
        asm volatile(NBODY_LT2(NBODY_SYM_PAIR) NBODY_SYM_PAIR(2) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 2 && MODE == 2)
        asm volatile(NBODY_LT2(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(2) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 3 && MODE == 0)
        asm volatile(NBODY_LT3(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 3 && MODE == 1)
        asm volatile(NBODY_LT3(NBODY_SYM_PAIR) NBODY_SYM_PAIR(3) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 3 && MODE == 2)
        asm volatile(NBODY_LT3(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(3) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 4 && MODE == 0)
        asm volatile(NBODY_LT4(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 4 && MODE == 1)
        asm volatile(NBODY_LT4(NBODY_SYM_PAIR) NBODY_SYM_PAIR(4) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 4 && MODE == 2)
        asm volatile(NBODY_LT4(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(4) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 5 && MODE == 0)
        asm volatile(NBODY_LT5(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 5 && MODE == 1)
        asm volatile(NBODY_LT5(NBODY_SYM_PAIR) NBODY_SYM_PAIR(5) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 5 && MODE == 2)
        asm volatile(NBODY_LT5(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(5) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 6 && MODE == 0)
        asm volatile(NBODY_LT6(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 6 && MODE == 1)
        asm volatile(NBODY_LT6(NBODY_SYM_PAIR) NBODY_SYM_PAIR(6) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 6 && MODE == 2)
        asm volatile(NBODY_LT6(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(6) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 7 && MODE == 0)
        asm volatile(NBODY_LT7(NBODY_SYM_PAIR) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 7 && MODE == 1)
        asm volatile(NBODY_LT7(NBODY_SYM_PAIR) NBODY_SYM_PAIR(7) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (!EQ && T == 7 && MODE == 2)
        asm volatile(NBODY_LT7(NBODY_SYM_PAIR) NBODY_SYM_PAIR_M(7) NBODY_ROR_GEN NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 0 && MODE == 0)
        asm volatile("s_nop 1\n" NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 0 && MODE == 1)
        asm volatile(NBODY_LT0(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(0) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 0 && MODE == 2)
        asm volatile(NBODY_LT0(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(0) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 1 && MODE == 0)
        asm volatile(NBODY_LT1(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 1 && MODE == 1)
        asm volatile(NBODY_LT1(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(1) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 1 && MODE == 2)
        asm volatile(NBODY_LT1(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(1) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 2 && MODE == 0)
        asm volatile(NBODY_LT2(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 2 && MODE == 1)
        asm volatile(NBODY_LT2(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(2) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 2 && MODE == 2)
        asm volatile(NBODY_LT2(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(2) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 3 && MODE == 0)
        asm volatile(NBODY_LT3(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 3 && MODE == 1)
        asm volatile(NBODY_LT3(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(3) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 3 && MODE == 2)
        asm volatile(NBODY_LT3(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(3) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 4 && MODE == 0)
        asm volatile(NBODY_LT4(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 4 && MODE == 1)
        asm volatile(NBODY_LT4(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(4) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 4 && MODE == 2)
        asm volatile(NBODY_LT4(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(4) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 5 && MODE == 0)
        asm volatile(NBODY_LT5(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 5 && MODE == 1)
        asm volatile(NBODY_LT5(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(5) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 5 && MODE == 2)
        asm volatile(NBODY_LT5(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(5) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 6 && MODE == 0)
        asm volatile(NBODY_LT6(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 6 && MODE == 1)
        asm volatile(NBODY_LT6(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(6) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 6 && MODE == 2)
        asm volatile(NBODY_LT6(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(6) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 7 && MODE == 0)
        asm volatile(NBODY_LT7(NBODY_SYM_PAIR_EQ) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 7 && MODE == 1)
        asm volatile(NBODY_LT7(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ(7) NBODY_ROR_EQ NBODY_DIAG_OPS);
    else if constexpr (EQ && T == 7 && MODE == 2)
        asm volatile(NBODY_LT7(NBODY_SYM_PAIR_EQ) NBODY_SYM_PAIR_EQ_M(7) NBODY_ROR_EQ NBODY_DIAG_OPS);
}

#undef NBODY_DIAG_OPS
#undef NBODY_LT0
#undef NBODY_LT1
#undef NBODY_LT2
#undef NBODY_LT3
#undef NBODY_LT4
#undef NBODY_LT5
#undef NBODY_LT6
#undef NBODY_LT7
#undef NBODY_ROR_GEN
#undef NBODY_ROR_EQ
#undef NBODY_SYM_PAIR_M
#undef NBODY_SYM_PAIR_EQ_M
#undef NBODY_SYM_PAIR_EQ_FIRST
#undef NBODY_SYM_PAIR_FIRST
#undef NBODY_SYM_BLOCK_ADD

#undef NBODY_SYM_PAIR_EQ
#undef NBODY_SYM_PAIR
#undef NBODY_ROR1

// The phases of block pair (P, Q): every wave streams the T tiles of block Q
// past its A i bodies per lane, adding the i side into acc and each tile's
// reaction into jacc. With EQ the pair weights leave out the (common) mass.
template <int A, int W, bool NEWTON, bool EQ, bool PIN>
__device__ __forceinline__ void sym_phases(const float4* __restrict__ pm, int n, float eps2, int Q,
                                           const Lanes<A>& me, const float (&mi)[A], Lanes<A>& acc,
                                           float (*jacc)[64 * A * W])
{
    constexpr int B = 64 * A * W;
    constexpr int T = A * W;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int ph = 0; ph < T; ++ph) {
        const int tile = (ph + wave * A) % T;
        const int jbase = Q * B + tile * 64;  // wave-uniform
        // a tile past the end of a ragged last block holds only zero-mass
        // padding, which would add exact zeros to the i side: one-wave
        // workgroups skip its work (the barrier is still met; with more waves
        // the extra branch costs registers, and those shapes serve large
        // systems where ragged blocks are rare)
        if (W > 1 || jbase < n) {
            const float4 b = lane < n - jbase ? pm[jbase + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            float xj = b.x, yj = b.y, zj = b.z, mj = b.w;
            float rx = 0.f, ry = 0.f, rz = 0.f;
            for (int k = 0; k < 64; ++k) {
                if constexpr (A == 8 && !NEWTON) {
                    if constexpr (EQ) sym_step8_eq<PIN>(xj, yj, zj, rx, ry, rz, acc, me, eps2);
                    else sym_step8<PIN>(xj, yj, zj, mj, rx, ry, rz, acc, me, mi, eps2);
                    continue;
                }
                if constexpr (A == 2 && !NEWTON) {
                    if constexpr (EQ) sym_step2_eq(xj, yj, zj, rx, ry, rz, acc, me, eps2);
                    else sym_step2(xj, yj, zj, mj, rx, ry, rz, acc, me, mi, eps2);
                    continue;
                }
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
                    const float si = EQ ? inv3 : mj * inv3;
                    const float sj = EQ ? -inv3 : -mi[u] * inv3;
                    acc.x[u] = __builtin_fmaf(dx, si, acc.x[u]);
                    acc.y[u] = __builtin_fmaf(dy, si, acc.y[u]);
                    acc.z[u] = __builtin_fmaf(dz, si, acc.z[u]);
                    rx = __builtin_fmaf(dx, sj, rx);
                    ry = __builtin_fmaf(dy, sj, ry);
                    rz = __builtin_fmaf(dz, sj, rz);
                }
                xj = wave_ror1(xj);
                yj = wave_ror1(yj);
                zj = wave_ror1(zj);
                if constexpr (!EQ) mj = wave_ror1(mj);
                rx = wave_ror1(rx);
                ry = wave_ror1(ry);
                rz = wave_ror1(rz);
            }
            const int slot = tile * 64 + lane;
            jacc[0][slot] += rx;
            jacc[1][slot] += ry;
            jacc[2][slot] += rz;
        }
        __syncthreads();
    }
}

// True on every thread when `ok` holds on every thread of the workgroup.
template <int W>
__device__ __forceinline__ bool workgroup_all(bool ok)
{
    if constexpr (W == 1) return __builtin_amdgcn_ballot_w64(!ok) == 0;
    else return __builtin_amdgcn_readfirstlane(__syncthreads_and(ok)) != 0;
}

// Block pair (P, Q), P < Q. `jacc` is the workgroup's LDS reaction buffer.
// P is never the last (possibly ragged) block, so every i lane holds a real
// body; j bodies past n enter with zero mass. When every body of both blocks
// has the mass of body P*B (uniform ICs: unit masses, m = 1/N), the pair
// weights leave the mass out and the partials are multiplied by it once, at
// the store; with m = 1 (or any power of two) the result is bit-identical to
// the general weights.
template <int A, int W, bool NEWTON, bool PIN = false>
__device__ __forceinline__ void sym_pair(const float4* __restrict__ pm, int n, float eps2, const SlotMap& slots, int P,
                                         int Q, float (*jacc)[64 * A * W])
{
    constexpr int B = 64 * A * W;
    const float m0 = pm[P * B].w;

    Lanes<A> me, acc;
    float mi[A];
    bool same = true;
#pragma unroll
    for (int u = 0; u < A; ++u) {
        const float4 q = pm[P * B + u * 64 * W + threadIdx.x];
        me.x[u] = q.x;
        me.y[u] = q.y;
        me.z[u] = q.z;
        mi[u] = q.w;
        same &= q.w == m0;
    }
    // all A loads in flight at once (no short-circuit), clamped into the array
#pragma unroll
    for (int u = 0; u < A; ++u) {
        const int j = Q * B + u * 64 * W + (int)threadIdx.x;
        same &= (j < n) & (pm[min(j, n - 1)].w == m0);
    }
    zero<A>(acc);

    for (int k = threadIdx.x; k < B; k += 64 * W) jacc[0][k] = jacc[1][k] = jacc[2][k] = 0.f;
    __syncthreads();
    const bool eq = workgroup_all<W>(same);
    if (eq) sym_phases<A, W, NEWTON, true, PIN>(pm, n, eps2, Q, me, mi, acc, jacc);
    else sym_phases<A, W, NEWTON, false, PIN>(pm, n, eps2, Q, me, mi, acc, jacc);
    const float scale = eq ? m0 : 1.0f;
    float* __restrict__ jdst = slots.j_side(P, Q, B);
    for (int k = threadIdx.x; k < B; k += 64 * W) {
        jdst[k] = jacc[0][k] * scale;
        jdst[slots.bplane + k] = jacc[1][k] * scale;
        jdst[2 * slots.bplane + k] = jacc[2][k] * scale;
    }
    float* __restrict__ idst = slots.i_side(P, Q, B);
    const size_t iplane = slots.i_plane(P);
#pragma unroll
    for (int u = 0; u < A; ++u) {
        const size_t o = u * 64 * W + threadIdx.x;
        idst[o] = acc.x[u] * scale;
        idst[iplane + o] = acc.y[u] * scale;
        idst[2 * iplane + o] = acc.z[u] * scale;
    }
}

// Block P against itself, each pair once, on one wave with 8 bodies per lane
// (blocks of 512). Tile t of the block rotates past the lanes as in a block
// pair and meets the i sub-blocks u < t in full, both directions. Against its
// own sub-block (u = t) it takes steps 1..32 only, step 32 on lanes 32..63:
// a pair of lanes at rotation distance d is met at step d and at step 64 - d,
// so this covers every pair of the 64 exactly once. When the tile's
// reactions come home after 64 steps they sit in the lanes that hold the
// same bodies' i sums, and are added there: no LDS, no cross-lane traffic.
template <int T, bool EQ>
__device__ __forceinline__ void sym_diag_phases(const float4* __restrict__ pm, int n, float eps2, int P,
                                                const Lanes<8>& me, const float (&mi)[8], Lanes<8>& acc)
{
    const int jbase = P * 512 + T * 64;  // wave-uniform
    if (jbase >= n) return;              // the rest of the block is padding
    const int lane = threadIdx.x & 63;
    const float4 b = lane < n - jbase ? pm[jbase + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    float xj = b.x, yj = b.y, zj = b.z, mj = b.w;
    float rx = 0.f, ry = 0.f, rz = 0.f;
    constexpr unsigned long long upper = 0xffffffff00000000ull;  // lanes 32..63
    // step 0: the tile meets itself lane by lane (self pairs, skipped)
    sym_diag_step8<T, 0, EQ>(xj, yj, zj, mj, rx, ry, rz, acc, me, mi, eps2, upper);
    for (int k = 1; k < 32; ++k) sym_diag_step8<T, 1, EQ>(xj, yj, zj, mj, rx, ry, rz, acc, me, mi, eps2, upper);
    sym_diag_step8<T, 2, EQ>(xj, yj, zj, mj, rx, ry, rz, acc, me, mi, eps2, upper);
    for (int k = 33; k < 64; ++k) sym_diag_step8<T, 0, EQ>(xj, yj, zj, mj, rx, ry, rz, acc, me, mi, eps2, upper);
    acc.x[T] += rx;
    acc.y[T] += ry;
    acc.z[T] += rz;
    if constexpr (T + 1 < 8) sym_diag_phases<T + 1, EQ>(pm, n, eps2, P, me, mi, acc);
}

// Block P's diagonal as one task (see sym_diag_phases), into slot P. Bodies
// past n are zero-mass padding and store into padding no reduction reads.
// (v_rsq_f32 without a Newton step, as the batched kernel uses it.)
__device__ __forceinline__ void sym_diag_block(const float4* __restrict__ pm, int n, int npad, float eps2,
                                               float* __restrict__ part, int P)
{
    constexpr int B = 512;
    const int lane = threadIdx.x & 63;
    const size_t plane = (size_t)(npad / B) * npad;
    const float m0 = pm[P * B].w;
    Lanes<8> me, acc;
    float mi[8];
    bool same = true;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int i = P * B + u * 64 + lane;
        const float4 q = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        me.x[u] = q.x;
        me.y[u] = q.y;
        me.z[u] = q.z;
        mi[u] = q.w;
        same &= (i < n) & (q.w == m0);
    }
    zero<8>(acc);
    const bool eq = __builtin_amdgcn_ballot_w64(!same) == 0;
    if (eq) sym_diag_phases<0, true>(pm, n, eps2, P, me, mi, acc);
    else sym_diag_phases<0, false>(pm, n, eps2, P, me, mi, acc);
    const float scale = eq ? m0 : 1.0f;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const size_t o = (size_t)P * npad + (size_t)P * B + u * 64 + lane;
        part[o] = acc.x[u] * scale;
        part[plane + o] = acc.y[u] * scale;
        part[2 * plane + o] = acc.z[u] * scale;
    }
}

// Piece k (of A) of block P's diagonal: its 64*W threads take bodies
// P*B + k*64*W + threadIdx.x against block P's bodies, a plain sum in j order,
// into slot P. Lanes past n compute on a zero body and store into padding
// that no reduction reads. CONST_VIEW reads the j bodies through the constant
// address space (pointers that come from memory rather than kernel arguments).
template <int B, bool NEWTON, bool CONST_VIEW>
__device__ __forceinline__ void sym_diag_piece(const float4* __restrict__ pm, int n, float eps2, const SlotMap& slots,
                                               int P, int k, int threads)
{
    if (P * B + k * threads >= n) return;  // a piece made only of padding (workgroup-uniform)
    const int i = P * B + k * threads + (int)threadIdx.x;
    const float4 q = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    Lanes<1> me, acc;
    me.x[0] = q.x;
    me.y[0] = q.y;
    me.z[0] = q.z;
    zero<1>(acc);
    const int j0 = P * B;
    if constexpr (CONST_VIEW) sweep_const<1, NEWTON>(pm + j0, min(n - j0, B), eps2, me, acc);
    else sweep_sgpr<1, NEWTON>(pm + j0, min(n - j0, B), eps2, me, acc);
    // slot P of block P's bodies: a band block, region B
    float* __restrict__ dst = slots.j_side(P, P, B) + k * threads + threadIdx.x;
    dst[0] = acc.x[0];
    dst[slots.bplane] = acc.y[0];
    dst[2 * slots.bplane] = acc.z[0];
}

}  // namespace nbody
