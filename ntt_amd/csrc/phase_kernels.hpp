// Two-pass radix-staged negacyclic NTT / INTT kernels for gfx950 (wave64).
//
// Reference: the two-phase structure of src/v04_montgomery.cu:196-565
// (NTTPhase1/NTTPhase2/INTTPhase1/INTTPhase2) and the config semantics of
// src/ntt_config.h:11-16.  The index formulation below is this library's own.
//
// A limb of N = 2^n words is viewed as a 2^R1 x 2^R2 row-major matrix
// (n = R1 + R2).  Forward stage s (0-based, Cooley-Tukey, half-distance
// 2^(n-1-s)) acts on bit b = n-1-s of the word index i and uses twiddle
// tw[2^(n-1-b) + (i >> (b+1))]; inverse (Gentleman-Sande) stages act on bits
// 0, 1, ..., n-1 in that order with the same index formula over the inverse
// table.  Bits >= R2 are row bits (column pass: butterflies run down each
// column), bits < R2 are in-row bits (row pass over contiguous rows).
//
// Inside a pass a thread owns E = 2^S words whose pass-local index v differs
// only in an S-bit window [lo, lo+S):
//     v = ((lane >> lo) << (lo + S)) | (x << lo) | (lane & (2^lo - 1)),
// x in [0, E).  One "round" performs every pending stage whose bit lies in the
// window, entirely in registers.  Between rounds the tile goes through LDS,
// kept in natural order: each round reads its window pattern and writes back
// to the same words, so one barrier per round suffices (no write-after-read
// hazard).  Twiddle index for the stage on window bit j:
//     (TB << (S-1-j)) + (x >> (j+1)),
//     TB = 2^(n - S - base - lo) + ext,  ext = (row bits above the window),
// where base is the position of the pass's lowest bit in i.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "arith.hpp"

namespace ntt {

// How the per-limb `scale` pointer is turned into the Montgomery multiplier
// applied on the first load of a forward transform's first pass, or on the
// last store of an inverse transform's last pass.
enum ScaleMode : int {
    SCALE_NONE = 0,          // no multiply
    SCALE_AS_IS = 1,         // multiply by scale[l] (already the Montgomery multiplier)
    SCALE_FROM_MONT = 2,     // multiplier = from_mont(scale[l]) (v04 degree_inv in Montgomery form)
    SCALE_TO_MONT = 3,       // multiplier = to_mont(scale[l]) (v03 plain degree_inv)
};

struct PhaseParams {
    uint32_t* A;              // operand base, operand k at A + k * limbs * N ...
    uint32_t* const* A_ptrs;  // ... or per-limb operand pointers (count == 1)
    const uint32_t* src;      // this pass's input, same layout as A (null: A)
    const uint32_t* sub;      // final forward pass, mod-down: store (sub - v) * post[l]
    const uint32_t* post;
    const uint32_t* tw;       // twiddle table, limb l at tw + l * N ...
    const uint32_t* const* tw_ptrs;  // ... or per-limb table pointers (twiddle cache)
    const uint32_t* moduli;   // q per limb
    const uint32_t* q_inv;    // -q^-1 mod 2^32 per limb, or null (derived)
    const uint32_t* scale;    // per-limb scale source, or null
    int scale_mode;           // ScaleMode
    int logN;                 // n
    int R;                    // stages in this pass (R1 for the column pass, R2 for the row pass)
    int W;                    // warp batching: log2 columns per tile / rows per workgroup
    uint32_t tiles;           // workgroups per limb (a power of two)
    uint32_t tiles_log2;
    // Division by count as a multiply-high: q = (t + ((x - t) >> s1)) >> s2
    // with t = mulhi(x, count_magic) (launch_pass sets the three fields).
    uint32_t count_magic, count_s1, count_s2;
    uint32_t limbs;           // limbs per operand
    uint32_t count;           // operands
    int xcd_remap;            // 1: group workgroups of the same limb index on one XCD
    int final_pass;           // 1: this pass ends the transform (outputs made canonical)
    // Limb mapping (contiguous operands only): the grid covers limbs - skip_len
    // limbs per operand, plan limbs [skip_start, skip_start + skip_len) are
    // skipped; the input of plan limb p of operand o is src (or A) +
    // (o * src_period + src_first + p) * N.  Mod-down: sub holds sub_end -
    // sub_start limbs per operand for plan limbs [sub_start, sub_end), each
    // optionally multiplied by subf[p - sub_start]; other limbs subtract from 0.
    uint32_t skip_start, skip_len;
    uint32_t src_period, src_first;
    uint32_t sub_start, sub_end;
    const uint32_t* subf;
};

// x / d for any 32-bit x, with (magic, s1, s2) from div_magic(d): the
// round-up method for division by an invariant integer (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994).
// With wave-uniform operands it runs on the scalar unit (s_mul_hi_u32)
// instead of a VALU reciprocal sequence.
__host__ __device__ __forceinline__ uint32_t div_by_magic(uint32_t x, uint32_t magic, uint32_t s1, uint32_t s2) {
    const uint32_t t = (uint32_t)(((uint64_t)x * magic) >> 32);
    return (t + ((x - t) >> s1)) >> s2;
}

// (magic, s1, s2) for divisor d >= 1: l = ceil(log2 d), magic = floor(2^32
// (2^l - d) / d) + 1, s1 = min(l, 1), s2 = max(l - 1, 0).
inline void div_magic(uint32_t d, uint32_t* magic, uint32_t* s1, uint32_t* s2) {
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    *magic = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    *s1 = l < 1 ? l : 1;
    *s2 = l > 1 ? l - 1 : 0;
}

// Arithmetic variants of the pass kernels.
enum PassMode : int {
    MODE_CANON = 0,    // canonical residues, q < 2^31, Montgomery-form table
    MODE_TWCONV = 1,   // canonical, normal-domain table converted on load (v03 drop-in)
    MODE_LAZY = 2,     // lazy residues, q < 2^30, Montgomery-form table
};

__device__ __forceinline__ uint32_t lds_pad(uint32_t i) { return i + (i >> 5); }

// Wave-local exchange fence: LDS operations of one wavefront complete in
// order, so ordering the compiler's memory operations is sufficient.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stages on window bits [jlo, jhi] (forward: from jhi down; inverse: from jlo up).
template <int S, bool INV, int MODE>
__device__ __forceinline__ void round_stages(uint32_t (&v)[1 << S], gcu32* __restrict__ tw,
                                             uint32_t TB, int jlo, int jhi, uint32_t q,
                                             uint32_t q_inv, uint32_t r2) {
    constexpr int E = 1 << S;
#pragma unroll
    for (int jj = 0; jj < S; ++jj) {
        const int j = INV ? jj : (S - 1 - jj);
        if (j < jlo || j > jhi) continue;
        const uint32_t base = TB << (S - 1 - j);
#pragma unroll
        for (int x = 0; x < E; ++x) {
            if (x & (1 << j)) continue;
            uint32_t w = tw[base + (x >> (j + 1))];
            if (MODE == MODE_TWCONV) w = mont_mul(w, r2, q, q_inv);
            if (MODE == MODE_LAZY) {
                if (INV)
                    gs_butterfly_lazy(v[x], v[x + (1 << j)], w, q, q_inv);
                else
                    ct_butterfly_lazy(v[x], v[x + (1 << j)], w, q, q_inv);
            } else if (INV) {
                gs_butterfly(v[x], v[x + (1 << j)], w, q, q_inv);
            } else {
                ct_butterfly(v[x], v[x + (1 << j)], w, q, q_inv);
            }
        }
    }
}

// One pass.  COL: column pass (outer stages; forward phase 1 / inverse phase 2).
// !COL: row pass (inner stages; forward phase 2 / inverse phase 1).
// RF > 0: the pass's radix stages fixed at compile time (RF == p.R), so the
// round windows and every index below fold to constants and shifts.  NF, WF
// > 0 (with RF): the transform size and the warp batching fixed as well
// (NF == p.logN, WF == p.W), so the tile geometry, the column stride and
// every LDS offset are constants too.
template <int S, bool INV, bool COL, int MODE, int RF = 0, int NF = 0, int WF = 0>
__global__ __launch_bounds__(1024) void phase_kernel(PhaseParams p) {
    constexpr int E = 1 << S;
    extern __shared__ uint32_t lds[];
    static_assert(!(NF || WF) || (NF && WF && RF), "NF and WF are set together, with RF");

    const int n = NF ? NF : p.logN, R = RF ? RF : p.R, W = WF ? WF : p.W;
    const uint32_t tiles = NF ? (1u << (NF - RF - WF)) : p.tiles;
    const uint32_t t = threadIdx.x;

    // ---- workgroup -> (limb index, operand, tile) ------------------------
    uint32_t b = blockIdx.x;
    if (p.xcd_remap) {
        const uint32_t B = gridDim.x;   // host guarantees B % 8 == 0
        b = (b & 7u) * (B >> 3) + (b >> 3);
    }
    const uint32_t tile = b & (tiles - 1);
    const uint32_t rest = b >> (NF ? (NF - RF - WF) : p.tiles_log2);
    const uint32_t lv = div_by_magic(rest, p.count_magic, p.count_s1, p.count_s2);
    const uint32_t op = rest - lv * p.count;
    const uint32_t li = lv + (lv >= p.skip_start ? p.skip_len : 0);   // plan limb
    const size_t N = (size_t)1 << n;
    const size_t off = ((size_t)op * p.limbs + li) * N;
    gu32* A = (gu32*)(p.A_ptrs ? p.A_ptrs[li] : p.A + off);
    gcu32* In = p.src ? (gcu32*)(p.src + ((size_t)op * p.src_period + p.src_first + li) * N) : A;
    gcu32* tw = (gcu32*)(p.tw_ptrs ? p.tw_ptrs[li] : p.tw + (size_t)li * N);

    const uint32_t q = p.moduli[li];
    const uint32_t q_inv = p.q_inv ? p.q_inv[li] : neg_inv_mod_2_32(q);
    uint32_t r2 = 0;
    if (MODE == MODE_TWCONV || p.scale_mode == SCALE_TO_MONT) r2 = r2_mod(q);
    uint32_t mult = 0;
    if (p.scale_mode != SCALE_NONE) {
        const uint32_t s = p.scale[li];
        mult = p.scale_mode == SCALE_AS_IS     ? s
             : p.scale_mode == SCALE_FROM_MONT ? mont_mul(s, 1u, q, q_inv)
                                               : mont_mul(s, r2, q, q_inv);
    }

    // ---- thread decomposition --------------------------------------------
    // COL: t = lane * 2^W + col, the tile is 2^W consecutive columns.
    // row: t = rib * 2^(R-S) + lane, the workgroup holds 2^W consecutive rows.
    const int lane_bits = R - S;
    uint32_t lane, col = 0, rib = 0, g = 0;
    if (COL) {
        col = t & ((1u << W) - 1);
        lane = t >> W;
        if (W >= 6) lane = __builtin_amdgcn_readfirstlane(lane);   // one lane-row per wave
    } else {
        lane = t & ((1u << lane_bits) - 1);
        rib = t >> lane_bits;
        g = (tile << W) + rib;
    }
    const uint32_t C = 1u << (n - R);                     // COL: columns per limb
    const uint32_t colbase = (tile << W) + col;
    const bool wave_local = !COL && lane_bits <= 6;       // a row never spans two waves

    auto vidx = [&](int lo, int x) -> uint32_t {
        return ((lane >> lo) << (lo + S)) | ((uint32_t)x << lo) | (lane & ((1u << lo) - 1));
    };
    auto gaddr = [&](uint32_t v) -> uint32_t {   // offset within the limb (N <= 2^24)
        return COL ? v * C + colbase : (g << R) + v;
    };
    auto laddr = [&](uint32_t v) -> uint32_t {
        return lds_pad(COL ? (v << W) + col : (rib << R) + v);
    };
    auto twbase = [&](int lo) -> uint32_t {
        const int base = COL ? (n - R) : 0;
        uint32_t ext = lane >> lo;
        if (!COL) ext |= g << (R - lo - S);
        return (1u << (n - S - base - lo)) + ext;
    };
    auto sync = [&]() {
        if (wave_local) wave_lds_sync(); else __syncthreads();
    };

    // Global tile I/O for a round with window `lo`; lo == 0 in a row pass
    // means every thread owns 2^S consecutive words (16-byte vector access).
    auto gload = [&](gcu32* G, int lo, uint32_t (&v)[E]) {
        if (!COL && lo == 0 && E >= 4) {
            gcu4* src = (gcu4*)(G + gaddr(vidx(0, 0)));
#pragma unroll
            for (int x = 0; x < E / 4; ++x) {
                const u32x4 w4 = src[x];
                v[4 * x] = w4.x; v[4 * x + 1] = w4.y; v[4 * x + 2] = w4.z; v[4 * x + 3] = w4.w;
            }
        } else {
#pragma unroll
            for (int x = 0; x < E; ++x) v[x] = G[gaddr(vidx(lo, x))];
        }
    };
    auto gstore = [&](int lo, const uint32_t (&v)[E]) {
        if (!COL && lo == 0 && E >= 4) {
            gu4* dst = (gu4*)(A + gaddr(vidx(0, 0)));
#pragma unroll
            for (int x = 0; x < E / 4; ++x)
                dst[x] = u32x4{v[4 * x], v[4 * x + 1], v[4 * x + 2], v[4 * x + 3]};
        } else {
#pragma unroll
            for (int x = 0; x < E; ++x) A[gaddr(vidx(lo, x))] = v[x];
        }
    };
    const bool scaled = p.scale_mode != SCALE_NONE;

    uint32_t v[E];
    const int K = (R + S - 1) / S;          // rounds
    const int rem = R - S * (K - 1);        // stages in the last round (1..S)
#pragma unroll

    // Forward: full rounds from the top window down, the partial round last
    // (window [0, S), stages on window bits rem-1 .. 0).  Inverse: full rounds
    // from the bottom window up, the partial round last (window [R-S, R),
    // stages on window bits S-rem .. S-1).
    for (int k = 0; k < K; ++k) {
        const bool last = (k == K - 1);
        const int lo = INV ? (last ? R - S : S * k) : (last ? 0 : R - S * (k + 1));
        if (k == 0) {
            gload(In, lo, v);
            if (!INV && scaled) {
#pragma unroll
                for (int x = 0; x < E; ++x)
                    v[x] = MODE == MODE_LAZY ? mont_mul_lazy(v[x], mult, q, q_inv) : mont_mul(v[x], mult, q, q_inv);
            }
        } else {
#pragma unroll
            for (int x = 0; x < E; ++x) v[x] = lds[laddr(vidx(lo, x))];
        }
        if (INV)
            round_stages<S, true, MODE>(v, tw, twbase(lo), last ? S - rem : 0, S - 1, q, q_inv, r2);
        else
            round_stages<S, false, MODE>(v, tw, twbase(lo), 0, last ? rem - 1 : S - 1, q, q_inv, r2);
        if (!last) {
#pragma unroll
            for (int x = 0; x < E; ++x) lds[laddr(vidx(lo, x))] = v[x];
            sync();
        } else {
            if (INV && scaled) {
#pragma unroll
                for (int x = 0; x < E; ++x) v[x] = mont_mul(v[x], mult, q, q_inv);
            } else if (MODE == MODE_LAZY && p.final_pass) {
                // forward values lie in [0, 4q), unscaled inverse ones in [0, 2q)
#pragma unroll
                for (int x = 0; x < E; ++x) v[x] = canon_4q(v[x], q);
            }
            if (!INV && !COL && p.sub) {
                // mod-down: (sub - NTT) * post, read at the positions stored below
                // (row passes only: the forward transform ends in its row pass)
                const uint32_t f = p.post[li];
                uint32_t s[E];
                if (li >= p.sub_start && li < p.sub_end) {
                    const uint32_t sl = li - p.sub_start;
                    gload((gcu32*)(p.sub + ((size_t)op * (p.sub_end - p.sub_start) + sl) * N), lo, s);
                    if (p.subf) {
                        const uint32_t g = p.subf[sl];
#pragma unroll
                        for (int x = 0; x < E; ++x) s[x] = mont_mul(s[x], g, q, q_inv);
                    }
                } else {
#pragma unroll
                    for (int x = 0; x < E; ++x) s[x] = 0;
                }
#pragma unroll
                for (int x = 0; x < E; ++x) v[x] = mont_mul(sub_mod(s[x], v[x], q), f, q, q_inv);
            }
            gstore(lo, v);
        }
    }
}

}  // namespace ntt
