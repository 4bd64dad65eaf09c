// k_pkmfma.hpp -- batched INT32 stage-1 scores on the matrix cores, read from
// the sliced packets (included by the k_pkmfma_*.hip instantiation units).
//
// The scores are exact integer dot products of the quantized rows with the
// int8 queries (quantify_util.py:5-6, 173-196), computed for a block of
// queries per pass over the matrix with v_mfma_i32_32x32x32_i8.  The matrix
// is read once per pass as its packets (about 2 bytes per non-zero); a
// workgroup expands one 64-row slice at a time into an LDS image with one
// byte per (row, column), and each wavefront holds the B operands of its
// query tiles in registers for the whole kernel.  Only the four best
// (score, row) keys of each (query, slice) leave the kernel.
//
// The wavefronts of a workgroup have two roles, one of each per SIMD.  MFMA
// wavefronts run a straight-line chain per slice: while it reads slice i's
// image, the same code reduces the accumulators of slice i-1 to their
// four-key lists.  Decode wavefronts meanwhile fill the other image with the
// background value and expand slice i+1's packets into it, with the packets
// of slice i+2 in flight.  So the matrix cores, the vector ALUs and the
// memory work at the same time.
#pragma once

#include "kernels_common.hpp"

#include <utility>

typedef int pv4i __attribute__((ext_vector_type(4)));
typedef int pv16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t pk_med3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

__device__ __forceinline__ int32_t pk_med3s(int32_t a, int32_t b, int32_t c) {
    int32_t d;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// keeps k0 >= k1 >= k2 >= k3 the four largest keys offered: the new i-th
// entry is the median of the old (i-1)-th, i-th and x
__device__ __forceinline__ void pk_top4(uint32_t &k0, uint32_t &k1, uint32_t &k2, uint32_t &k3, uint32_t x) {
    const uint32_t n3 = pk_med3(k2, k3, x), n2 = pk_med3(k1, k2, x), n1 = pk_med3(k0, k1, x);
    k0 = max(k0, x);
    k1 = n1;
    k2 = n2;
    k3 = n3;
}

// The same over signed keys.
__device__ __forceinline__ void pk_top4s(int32_t &k0, int32_t &k1, int32_t &k2, int32_t &k3, int32_t x) {
    const int32_t n3 = pk_med3s(k2, k3, x), n2 = pk_med3s(k1, k2, x), n1 = pk_med3s(k0, k1, x);
    k0 = max(k0, x);
    k1 = n1;
    k2 = n2;
    k3 = n3;
}

// Pre-key of accumulator value rr of row tile rt (a lane's rows rt*32 +
// (rr/4)*8 + rr%4, plus 4 in the upper half-wave): (score << 6) | row, signed
// (|score| < 2^25); the row part is an inline constant.
__device__ __forceinline__ int32_t pk_prekey(int32_t v, int rt, int rr) {
    return (int32_t)(((uint32_t)v << 6) | (uint32_t)(rt * 32 + (rr >> 2) * 8 + (rr & 3)));
}

// Workgroup barrier for LDS only: waits for this wavefront's LDS operations,
// not for its global loads (a full barrier would also wait for the packet
// loads issued for the slices ahead).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Value modes of the image bytes: the value field as is (unsigned below 8
// bits, and signed 8-bit whose field is the two's-complement byte), unsigned
// 8-bit stored as v - 128 (the image background is -128 and 128 * sum(query)
// is added back per query), and signed below 8 bits (sign-extended).
// Entry j of a lane's image-packet chunk (E dwords = 32 entries) stored into
// the decode wavefront's 16 image rows at wimg: an entry is byte | (o << 8)
// with o = row * stride + column inside those rows, and its byte is already
// the image's value.
template <int E>
__device__ __forceinline__ void pk_entry(const uint32_t (&w)[E], int j, unsigned char *wimg, int ob) {
    const int pos = j * E, wi = pos >> 5, sh = pos & 31;
    const uint32_t hi = (wi + 1 < E) ? w[wi + 1] : 0u;
    const uint32_t bits = (sh + E <= 32) ? (w[wi] >> sh) : __builtin_amdgcn_alignbit(hi, w[wi], sh);
    wimg[__builtin_amdgcn_ubfe(bits, 8, ob)] = (unsigned char)bits;
}

// Wavefront roles: up to kPkMfmaWaves MFMA wavefronts (one or two 32-query
// tiles each, their B operands in registers) and kPkDecodeWaves decode
// wavefronts (packets into the LDS images), one of each per SIMD: while one
// SIMD's MFMA wavefront multiplies, its decode wavefront expands the next
// slice.
struct PkShape {
    // MODE 3: eight MFMA wavefronts (two per SIMD) share the decode of each
    // slice, three wavefronts per SIMD
    static constexpr int threads(int mode) { return (mode == 3 ? 12 : kPkMfmaWaves + kPkDecodeWaves) * 64; }
    // two wavefronts per SIMD; four (two workgroups per CU) for the
    // one-tile half-round chain, whose registers fit 128
    static constexpr int waves_per_eu(int mode) { return mode == 1 ? 4 : mode == 3 ? 3 : 2; }
};

// Registers of an MFMA wavefront (one object, so that the compile-time steps
// below reach them; every index is a constant).  Rounds alternate between
// parity 0 and 1: round P accumulates slice s into acc[P] while it reduces
// acc[1-P] (slice s - 1) to the four-key lists.
struct PkMRegs {
    pv4i ar[4][2];    // A operands in flight (ring)
    pv16i acc[2][2];  // [round parity][row tile]
    int32_t kk[4];    // four-pre-key list of the previous slice
};

// What every round of pk_scores uses.
struct PkCtx {
    const DenseArgs *a;
    unsigned char *smem;
    int lane, hi, wave;
    int64_t end;  // one past this workgroup's last slice
    uint32_t cols_pad, stride, img_bytes, fill, a_base;
    uint32_t wrow;                // decode wavefronts: offset of their 16 rows in an image (16 * dw * stride)
    int ob;                       // offset bits of the image-packet entries
    __amdgpu_buffer_rsrc_t meta;  // a.pslices
};

// Step kb of an MFMA round of parity P and the list work scheduled with it:
// a compile-time step, so the B operands and accumulators are addressed by
// constants.
template <int KB, int P, bool LISTS, int kb>
__device__ __forceinline__ void pk_mstep(const PkCtx &c, PkMRegs &r, const pv4i (&B)[KB], const pv4i *A0,
                                         const pv4i *A1) {
    constexpr int RING = 4, DIST = 2;  // A operands: ring slots, column blocks ahead
    if constexpr (kb + DIST < KB) {    // constant offsets; blocks past nkb meet zero B operands
        r.ar[(kb + DIST) % RING][0] = A0[(kb + DIST) * 2];
        r.ar[(kb + DIST) % RING][1] = A1[(kb + DIST) * 2];
    }
    __builtin_amdgcn_sched_barrier(0);
    r.acc[P][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(r.ar[kb % RING][0], B[kb], r.acc[P][0], 0, 0, 0);
    r.acc[P][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(r.ar[kb % RING][1], B[kb], r.acc[P][1], 0, 0, 0);
    // the previous slice's accumulators into the list, value v at step v*KB/32
#pragma unroll
    for (int v = 0; v < 32; ++v)
        if (LISTS && v * KB / 32 == kb)
            pk_top4s(r.kk[0], r.kk[1], r.kk[2], r.kk[3], pk_prekey(r.acc[1 - P][v / 16][v & 15], v / 16, v & 15));
    __builtin_amdgcn_sched_barrier(0);
}

template <int KB, int P, bool LISTS, int... KBS>
__device__ __forceinline__ void pk_mchain(std::integer_sequence<int, KBS...>, const PkCtx &c, PkMRegs &r,
                                          const pv4i (&B)[KB], const pv4i *A0, const pv4i *A1) {
    (pk_mstep<KB, P, LISTS, KBS>(c, r, B, A0, A1), ...);
}

// Lanes 0..31 receive the value of lane + 32 (v_permlane32_swap; lanes
// 32..63 keep their own).
__device__ __forceinline__ uint32_t pk_upper_half(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_permlane32_swap((int)v, (int)v, false, false)[1];
}

// The four-key list of this wavefront's tile (each lane's pre-keys k0..k3,
// taken modulo 2^32): offset to the query's keys ((score + corr + 2^25) << 6
// | row, unsigned), both halves of the lane pair merged, stored for slice sx.
__device__ __forceinline__ void pk_store_list(const PkCtx &c, uint32_t Dk, int q, uint32_t k0, uint32_t k1,
                                              uint32_t k2, uint32_t k3, int64_t sx) {
    k0 += Dk;
    k1 += Dk;
    k2 += Dk;
    k3 += Dk;
    // the upper half-wave's lists into lanes 0..31 (the lanes that store)
    const uint32_t o0 = pk_upper_half(k0), o1 = pk_upper_half(k1), o2 = pk_upper_half(k2), o3 = pk_upper_half(k3);
    pk_top4(k0, k1, k2, k3, o0);
    pk_top4(k0, k1, k2, k3, o1);
    pk_top4(k0, k1, k2, k3, o2);
    pk_top4(k0, k1, k2, k3, o3);
    const DenseArgs &a = *c.a;
    if (c.hi == 0 && q < a.nq) *(uint4 *)(a.tk + ((size_t)q * a.nslices + sx) * 4) = make_uint4(k0, k1, k2, k3);
}

// MFMA round of parity P: the slice in image `img` multiplied into acc[P];
// the list of the previous slice (acc[1-P], none for the first slice:
// !LISTS) reduced during the chain and stored.
template <int KB, int P, bool LISTS>
__device__ __forceinline__ void pk_mround(const PkCtx &c, PkMRegs &r, const pv4i (&B)[KB], uint32_t Dk, int q,
                                          int64_t s_prev, uint32_t img) {
    const unsigned char *cur = c.smem + img * c.img_bytes;
    r.acc[P][0] = (pv16i){0};
    r.acc[P][1] = (pv16i){0};
#pragma unroll
    for (int e = 0; e < 4; ++e) r.kk[e] = INT32_MIN;
    const pv4i *A0 = (const pv4i *)(cur + c.a_base);
    const pv4i *A1 = (const pv4i *)(cur + c.a_base + 32u * c.stride);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        r.ar[j][0] = A0[j * 2];
        r.ar[j][1] = A1[j * 2];
    }
    pk_mchain<KB, P, LISTS>(std::make_integer_sequence<int, KB>(), c, r, B, A0, A1);
    if constexpr (LISTS) pk_store_list(c, Dk, q, r.kk[0], r.kk[1], r.kk[2], r.kk[3], s_prev);
    lds_barrier();  // this slice's image read, the next one written
}

// The list of the last slice (accumulators of parity P).
template <int P>
__device__ __forceinline__ void pk_last_list(const PkCtx &c, const PkMRegs &r, uint32_t Dk, int q, int64_t sx) {
    int32_t k0 = INT32_MIN, k1 = INT32_MIN, k2 = INT32_MIN, k3 = INT32_MIN;
#pragma unroll
    for (int v = 0; v < 32; ++v) pk_top4s(k0, k1, k2, k3, pk_prekey(r.acc[P][v / 16][v & 15], v / 16, v & 15));
    pk_store_list(c, Dk, q, k0, k1, k2, k3, sx);
}

// An MFMA wavefront: query tile `tile` of block qb over slices s .. end-1.
template <int KB>
__device__ __forceinline__ void pk_mmain(const PkCtx &c, int qb, int tile, int64_t s) {
    const DenseArgs &a = *c.a;
    // B operands of the tile (zero past nkb)
    pv4i B[KB];
    {
        const pv4i *q = (const pv4i *)a.qtab + (size_t)qb * a.nkb * a.NT * 64 + c.lane;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            B[kb] = kb < a.nkb ? q[((size_t)kb * a.NT + tile) * 64] : (pv4i){0, 0, 0, 0};
    }
    // the lane's query and its key offset D = ((corr + 2^25) << 6) + 4*hi
    const int q = qb * a.NT * 32 + tile * 32 + (c.lane & 31);
    const uint32_t Dk = ((uint32_t)(a.corr[q] + (1 << 25)) << 6) + (uint32_t)(c.hi * 4);
    PkMRegs r;
    lds_barrier();  // the images filled
    lds_barrier();  // the first slice decoded
    // slice i of the range lies in image i % 2
    uint32_t img = 0;
    auto next_img = [&]() { img ^= 1u; };
    pk_mround<KB, 0, false>(c, r, B, Dk, q, -1, img);
    int64_t s_prev = s;
    ++s;
    for (;;) {
        if (s >= c.end) {
            pk_last_list<0>(c, r, Dk, q, s_prev);
            return;
        }
        next_img();
        pk_mround<KB, 1, true>(c, r, B, Dk, q, s_prev, img);
        s_prev = s++;
        if (s >= c.end) {
            pk_last_list<1>(c, r, Dk, q, s_prev);
            return;
        }
        next_img();
        pk_mround<KB, 0, true>(c, r, B, Dk, q, s_prev, img);
        s_prev = s++;
    }
}

// Registers of an MFMA wavefront that multiplies a slice in two
// half-rounds, one per row tile (32 rows), for NTL query tiles: every A read
// feeds NTL MFMAs, and only one row tile's accumulators are live per
// half-round besides the other row tile's being reduced.  NTL = 2 halves the
// LDS A reads per query (both tiles' B operands fit the registers at up to
// 16 column blocks); NTL = 1 fits 128 registers at up to 16 column blocks,
// so that two workgroups share a CU.  Half-round rt accumulates row tile rt
// into acc[rt] while it reduces acc[1 - rt] (the other row tile, from the
// half-round before) into the lists kk.
template <int NTL>
struct PkMRegsH {
    pv4i ar[4];          // A operands in flight (ring)
    pv16i acc[2][NTL];   // [row tile][query tile]
    int32_t kk[NTL][4];  // [query tile]: four-pre-key list of the slice being reduced
};

template <int KB, int NTL, int RT, bool LISTS, int kb>
__device__ __forceinline__ void pk_hstep(PkMRegsH<NTL> &r, const pv4i (&B)[NTL][KB], const pv4i *A) {
    constexpr int RING = 4, DIST = 2;
    if constexpr (kb + DIST < KB) r.ar[(kb + DIST) % RING] = A[(kb + DIST) * 2];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < NTL; ++t)
        r.acc[RT][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(r.ar[kb % RING], B[t][kb], r.acc[RT][t], 0, 0, 0);
    // the other row tile's accumulators into the lists, value v at step v*KB/(16*NTL)
#pragma unroll
    for (int v = 0; v < 16 * NTL; ++v)
        if (LISTS && v * KB / (16 * NTL) == kb)
            pk_top4s(r.kk[v / 16][0], r.kk[v / 16][1], r.kk[v / 16][2], r.kk[v / 16][3],
                     pk_prekey(r.acc[1 - RT][v / 16][v & 15], 1 - RT, v & 15));
    __builtin_amdgcn_sched_barrier(0);
}

template <int KB, int NTL, int RT, bool LISTS, int... KBS>
__device__ __forceinline__ void pk_hchain(std::integer_sequence<int, KBS...>, PkMRegsH<NTL> &r,
                                          const pv4i (&B)[NTL][KB], const pv4i *A) {
    (pk_hstep<KB, NTL, RT, LISTS, KBS>(r, B, A), ...);
}

// Half-round RT: row tile RT of the slice in image `img` into acc[RT].  Row
// tile 0 completes the previous slice's lists (its row tile 1, when LISTS);
// row tile 1 starts this slice's lists (its row tile 0).
template <int KB, int NTL, int RT, bool LISTS>
__device__ __forceinline__ void pk_hround(const PkCtx &c, PkMRegsH<NTL> &r, const pv4i (&B)[NTL][KB], uint32_t img) {
#pragma unroll
    for (int t = 0; t < NTL; ++t) r.acc[RT][t] = (pv16i){0};
    if constexpr (RT == 1) {
#pragma unroll
        for (int t = 0; t < NTL; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) r.kk[t][e] = INT32_MIN;
    }
    const pv4i *A = (const pv4i *)(c.smem + img * c.img_bytes + c.a_base + (uint32_t)RT * 32u * c.stride);
    r.ar[0] = A[0];
    r.ar[1] = A[2];
    pk_hchain<KB, NTL, RT, LISTS>(std::make_integer_sequence<int, KB>(), r, B, A);
}

// An MFMA wavefront with query tiles t0 .. t0 + NTL-1 of block qb over
// slices s .. end-1, in half-rounds (the same barriers as pk_mmain: two
// before the first slice, one after each slice's second half-round).
template <int KB, int NTL>
__device__ __forceinline__ void pk_mmainh(const PkCtx &c, int qb, int t0, int64_t s) {
    const DenseArgs &a = *c.a;
    pv4i B[NTL][KB];
    {
        const pv4i *q = (const pv4i *)a.qtab + (size_t)qb * a.nkb * a.NT * 64 + c.lane;
#pragma unroll
        for (int t = 0; t < NTL; ++t)
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
                B[t][kb] = kb < a.nkb ? q[((size_t)kb * a.NT + t0 + t) * 64] : (pv4i){0, 0, 0, 0};
    }
    int q[NTL];
    uint32_t Dk[NTL];
#pragma unroll
    for (int t = 0; t < NTL; ++t) {
        q[t] = qb * a.NT * 32 + (t0 + t) * 32 + (c.lane & 31);
        Dk[t] = ((uint32_t)(a.corr[q[t]] + (1 << 25)) << 6) + (uint32_t)(c.hi * 4);
    }
    PkMRegsH<NTL> r;
    lds_barrier();  // the images filled
    lds_barrier();  // the first slice decoded
    uint32_t img = 0;  // slice i of the range lies in image i % 2
    pk_hround<KB, NTL, 0, false>(c, r, B, img);
    pk_hround<KB, NTL, 1, true>(c, r, B, img);
    lds_barrier();  // this slice's image read, the next one written
    int64_t s_prev = s++;
    for (; s < c.end; s_prev = s++) {
        img ^= 1u;
        pk_hround<KB, NTL, 0, true>(c, r, B, img);
#pragma unroll
        for (int t = 0; t < NTL; ++t)
            pk_store_list(c, Dk[t], q[t], r.kk[t][0], r.kk[t][1], r.kk[t][2], r.kk[t][3], s_prev);
        pk_hround<KB, NTL, 1, true>(c, r, B, img);
        lds_barrier();
    }
    // the last slice's row tile 1
#pragma unroll
    for (int v = 0; v < 16 * NTL; ++v)
        pk_top4s(r.kk[v / 16][0], r.kk[v / 16][1], r.kk[v / 16][2], r.kk[v / 16][3],
                 pk_prekey(r.acc[1][v / 16][v & 15], 1, v & 15));
#pragma unroll
    for (int t = 0; t < NTL; ++t)
        pk_store_list(c, Dk[t], q[t], r.kk[t][0], r.kk[t][1], r.kk[t][2], r.kk[t][3], s_prev);
}

// The chain of MODE 3 (two MFMA wavefronts per SIMD, one query tile each):
// half-rounds with one accumulator set, the row tile reduced to the list
// after its chain -- the other MFMA wavefront of the SIMD multiplies
// meanwhile -- and an A ring of three slots, so that the B operands of up to
// 30 column blocks fit a third of the SIMD's registers.
template <int KB, int kb>
__device__ __forceinline__ void pk_sstep(pv4i (&ar)[3], pv16i &acc, const pv4i (&B)[KB], const pv4i *A) {
    constexpr int RING = 3, DIST = 2;
    if constexpr (kb + DIST < KB) ar[(kb + DIST) % RING] = A[(kb + DIST) * 2];
    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[kb % RING], B[kb], acc, 0, 0, 0);
}

template <int KB, int... KBS>
__device__ __forceinline__ void pk_schain(std::integer_sequence<int, KBS...>, pv4i (&ar)[3], pv16i &acc,
                                          const pv4i (&B)[KB], const pv4i *A) {
    (pk_sstep<KB, KBS>(ar, acc, B, A), ...);
}

// Row tile RT of the slice in image img into the list k0..k3.
template <int KB, int RT>
__device__ __forceinline__ void pk_shalf(const PkCtx &c, const pv4i (&B)[KB], uint32_t img, int32_t &k0, int32_t &k1,
                                         int32_t &k2, int32_t &k3) {
    const pv4i *A = (const pv4i *)(c.smem + img * c.img_bytes + c.a_base + (uint32_t)RT * 32u * c.stride);
    pv4i ar[3];
    ar[0] = A[0];
    ar[1] = A[2];
    pv16i acc = (pv16i){0};
    pk_schain<KB>(std::make_integer_sequence<int, KB>(), ar, acc, B, A);
#pragma unroll
    for (int v = 0; v < 16; ++v) pk_top4s(k0, k1, k2, k3, pk_prekey(acc[v], RT, v));
}

// An MFMA wavefront of MODE 3: query tile `tile` of block qb over slices
// s .. end-1 (the same barriers as pk_mmain).
template <int KB>
__device__ __forceinline__ void pk_mmains(const PkCtx &c, int qb, int tile, int64_t s) {
    const DenseArgs &a = *c.a;
    pv4i B[KB];
    {
        const pv4i *q = (const pv4i *)a.qtab + (size_t)qb * a.nkb * a.NT * 64 + c.lane;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            B[kb] = kb < a.nkb ? q[((size_t)kb * a.NT + tile) * 64] : (pv4i){0, 0, 0, 0};
    }
    const int q = qb * a.NT * 32 + tile * 32 + (c.lane & 31);
    const uint32_t Dk = ((uint32_t)(a.corr[q] + (1 << 25)) << 6) + (uint32_t)(c.hi * 4);
    lds_barrier();  // the images filled
    lds_barrier();  // the first slice decoded
    uint32_t img = 0;  // slice i of the range lies in image i % 2
    for (;;) {
        int32_t k0 = INT32_MIN, k1 = INT32_MIN, k2 = INT32_MIN, k3 = INT32_MIN;
        pk_shalf<KB, 0>(c, B, img, k0, k1, k2, k3);
        pk_shalf<KB, 1>(c, B, img, k0, k1, k2, k3);
        pk_store_list(c, Dk, q, k0, k1, k2, k3, s);
        lds_barrier();  // this slice's image read, the next one written
        if (++s >= c.end) return;
        img ^= 1u;
    }
}

// Decode wavefront dw owns rows 16*dw .. 16*dw + 15 of every image.  Its
// stream of a slice holds the entries of those rows split among its lanes,
// consecutive ones per lane, n per lane (n a multiple of 4, the same for
// every lane; a lane's last slots repeat entries of the next lanes), so the
// entries are spread evenly over the lanes whatever the rows' lengths.  Every byte of those rows is written
// by this wavefront only, in program order, so clearing an image and
// expanding the next slice into it need no barrier between them.
template <int E>
struct PkDRegs {
    uint32_t w[2][E];       // the first chunk of the next slice, by round parity
};

// This wavefront's stream of the slice with metadata m (offset in units of
// 64 dwords; n/4 of wavefront w in byte w): a buffer over its dwords (loads
// past its end return zeros and fetch nothing) and n.
struct PkStream {
    __amdgpu_buffer_rsrc_t rs;
    int n;
};

template <int E>
__device__ __forceinline__ PkStream pk_stream(const PkCtx &c, uint2 m, int dw) {
    const uint32_t off = __builtin_amdgcn_readfirstlane(m.x), cnt = __builtin_amdgcn_readfirstlane(m.y);
    dw = __builtin_amdgcn_readfirstlane(dw);  // uniform: the buffer descriptor stays scalar
    uint32_t woff = 0;
    for (int w = 0; w < dw; ++w) woff += (((cnt >> (8 * w)) & 255u) * 4u * E + 31u) / 32u;
    const uint32_t n = ((cnt >> (8 * dw)) & 255u) * 4u;
    PkStream st;
    st.rs = __builtin_amdgcn_make_buffer_rsrc((void *)(c.a->pwords + ((size_t)off + woff) * 64), (short)0,
                                              (int)((n * E + 31u) / 32u * 256u), 0x00020000 /* 32-bit data format */);
    st.n = (int)n;
    return st;
}

// The E dwords of chunk k (entries 32k .. 32k+31) of the lane's stream.
template <int E>
__device__ __forceinline__ void pk_load_chunk(uint32_t (&w)[E], __amdgpu_buffer_rsrc_t rs, const PkCtx &c, int k) {
#pragma unroll
    for (int i = 0; i < E; ++i)
        w[i] = __builtin_amdgcn_raw_buffer_load_b32(rs, ((k * E + i) * 64 + c.lane) * 4, 0, 2 /* nt */);
}

// This wavefront's rows of image img back to the background: 16 rows =
// stride 16-byte chunks (fewer LDS cycles than clearing the bytes the last
// slice wrote, one at a time, at every width up to 1024 columns).
__device__ __forceinline__ void pk_clear(const PkCtx &c, unsigned char *img) {
    pv4i *p4 = (pv4i *)(img + c.wrow);
    const pv4i f = {(int)c.fill, (int)c.fill, (int)c.fill, (int)c.fill};
    uint32_t i = (uint32_t)c.lane;
    for (; i + 192u < c.stride; i += 256u) {
        p4[i] = f;
        p4[i + 64u] = f;
        p4[i + 128u] = f;
        p4[i + 192u] = f;
    }
    for (; i < c.stride; i += 64u) p4[i] = f;
}

// The first m entries of a chunk (m wave-uniform, a multiple of 4).
template <int E>
__device__ __forceinline__ void pk_chunk(const uint32_t (&w)[E], int m, unsigned char *wimg, int ob) {
    m = __builtin_amdgcn_readfirstlane(m);  // a scalar branch per four entries
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (4 * q >= m) return;
#pragma unroll
        for (int j = 4 * q; j < 4 * q + 4; ++j) pk_entry<E>(w, j, wimg, ob);
    }
}

// A slice (stream st, first chunk in w, further chunks loaded here) into
// image img.
template <int E>
__device__ __forceinline__ void pk_expand(const PkCtx &c, const uint32_t (&w)[E], const PkStream &st,
                                          unsigned char *img) {
    unsigned char *wimg = img + c.wrow;
    pk_chunk<E>(w, st.n, wimg, c.ob);
    for (int k = 1; 32 * k < st.n; ++k) {
        uint32_t x[E];
        pk_load_chunk<E>(x, st.rs, c, k);
        pk_chunk<E>(x, st.n - 32 * k, wimg, c.ob);
    }
}

// Metadata of slice min(s, end - 1) as a vector load: a scalar load would be
// waited for at the next LDS barrier (both count in lgkmcnt), exposing its
// latency once per round.
__device__ __forceinline__ uint2 pk_slice_meta(const PkCtx &c, int64_t s) {
    const uint32_t i = (uint32_t)min(s, c.end - 1);
    uint2 m;
    m.x = __builtin_amdgcn_raw_buffer_load_b32(c.meta, (int)(i * 8u), 0, 0);
    m.y = __builtin_amdgcn_raw_buffer_load_b32(c.meta, (int)(i * 8u + 4u), 0, 0);
    return m;
}

// Decode round of parity P (the MFMA wavefronts multiply slice s from image
// P): slice s + 1 (metadata sl1, first chunk in r.w[P], requested a round
// earlier) expanded into image 1 - P after this wavefront's rows of it are
// cleared; the first chunk of slice s + 2 (sl2) requested into r.w[1-P].
// sl1, sl2 advance by one slice.
template <int E, int P>
__device__ __forceinline__ void pk_dround(const PkCtx &c, PkDRegs<E> &r, int dw, int64_t s, uint2 &sl1, uint2 &sl2) {
    unsigned char *nxt = c.smem + (1 - P) * c.img_bytes;
    const int64_t s3 = s + 3;
    const uint2 s3m = pk_slice_meta(c, s3);
    pk_load_chunk<E>(r.w[1 - P], pk_stream<E>(c, sl2, dw).rs, c, 0);
    pk_clear(c, nxt);
    pk_expand<E>(c, r.w[P], pk_stream<E>(c, sl1, dw), nxt);
    sl1 = sl2;
    sl2 = s3 < c.end ? s3m : make_uint2(0u, 0u);
    lds_barrier();  // the next slice's image complete
}

// A decode wavefront (dw = 0 .. kPkDecodeWaves-1) over slices s .. end-1
// (slices past the range read as empty).  Slice i of the range lies in image
// i % 2.  RAISE: with one MFMA wavefront per SIMD the decode wavefronts bound
// each round (the MFMA wavefronts wait at its barrier), so their
// instructions, mostly LDS stores, issue first; with two MFMA wavefronts per
// SIMD these bound it, and the decode keeps the default priority.
template <int E, bool RAISE>
__device__ __forceinline__ void pk_dmain(const PkCtx &c, int dw, int64_t s) {
    PkDRegs<E> r;
    if constexpr (RAISE) __builtin_amdgcn_s_setprio(2);
    pk_clear(c, c.smem);  // image 0's rows of this wavefront hold the background
    const uint2 s0m = pk_slice_meta(c, s);
    const uint2 s1m = pk_slice_meta(c, s + 1), s2m = pk_slice_meta(c, s + 2);
    uint2 sl1 = s + 1 < c.end ? s1m : make_uint2(0u, 0u);
    uint2 sl2 = s + 2 < c.end ? s2m : make_uint2(0u, 0u);
    {
        uint32_t w0[E];
        const PkStream st0 = pk_stream<E>(c, s0m, dw);
        pk_load_chunk<E>(w0, st0.rs, c, 0);
        pk_load_chunk<E>(r.w[0], pk_stream<E>(c, sl1, dw).rs, c, 0);
        lds_barrier();  // (the MFMA wavefronts' first barrier)
        pk_expand<E>(c, w0, st0, c.smem);
    }
    lds_barrier();  // the first slice expanded
    pk_dround<E, 0>(c, r, dw, s, sl1, sl2);
    ++s;
    for (;;) {
        if (s >= c.end) return;
        pk_dround<E, 1>(c, r, dw, s, sl1, sl2);
        if (++s >= c.end) return;
        pk_dround<E, 0>(c, r, dw, s, sl1, sl2);
        ++s;
    }
}

// Scores of the 64-row slices against query block qb (QB = 32 * NT queries;
// MFMA wavefronts of one tile each, or two, then kPkDecodeWaves decode
// wavefronts).  MODE 0: one tile per MFMA wavefront, whole-slice rounds; 1:
// one tile, half-rounds, registers for two workgroups per CU (up to 16
// column blocks); 2: two tiles, half-rounds (up to 16 column blocks).
// Grid = nwx * nqb workgroups; the nqb blocks of one slice range
// wx get workgroup ids 8 apart, hence one XCD, and the second and later
// blocks find the packets in its L2.  Workgroup wx takes the slices wx*per
// .. wx*per + per-1 (per = ceil(nslices / nwx)): consecutive slices, so that
// a query's four-key lists of one workgroup fill whole cache lines in L2
// before they are written back.  KB: column blocks the B registers hold (nkb <= KB; the chain runs over
// all KB with zero B operands past nkb, so that it has no control flow).
// Every wavefront passes the same barriers: two before the first round, then
// one per round (its end: the next slice's image complete, this one's read).
// Two images: the MFMA wavefronts multiply slice s from image s % 2 while the
// decode wavefronts clear and refill image (s + 1) % 2.
//
// Image of a slice (LDS): row r's bytes at r * stride + column, stride =
// cols_pad + 16; the bytes past cols meet zero B operands.  The A
// operand (row tile rt, column block kb) of lane l is the 16 bytes of row
// rt*32 + l%32 at column kb*32 + (l/32)*16: with stride/4 = 8m + 4, any 8
// consecutive lanes read disjoint banks.  The query table pairs each B byte
// with the same column.
template <int E, int KB, int MODE>
__global__ __launch_bounds__(PkShape::threads(MODE))
__attribute__((amdgpu_waves_per_eu(PkShape::waves_per_eu(MODE), PkShape::waves_per_eu(MODE)))) void pk_scores(
    DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = (int)(threadIdx.x >> 6), lane = lane_id(), hi = lane >> 5;
    const int nqb = a.nqb;
    const int c8 = (int)blockIdx.x % (8 * nqb);
    const int qb = c8 / 8;
    const int wx = ((int)blockIdx.x / (8 * nqb)) * 8 + (c8 % 8), nwx = (int)gridDim.x / nqb;
    const int64_t per = (a.nslices + nwx - 1) / nwx, s0 = (int64_t)wx * per;
    if (s0 >= a.nslices) return;  // the whole workgroup
    PkCtx c;
    c.a = &a;
    c.smem = smem;
    c.lane = lane;
    c.hi = hi;
    c.end = min(s0 + per, a.nslices);
    c.wave = wave;
    c.cols_pad = (uint32_t)a.cols_pad;
    c.stride = c.cols_pad + 16u;
    c.img_bytes = 64u * c.stride;
    c.fill = a.u8off ? 0x80808080u : 0u;                               // the image byte of an absent entry
    const int mw = a.NT / (MODE == 2 ? 2 : 1);                         // MFMA wavefronts
    c.wrow = (uint32_t)(wave - mw) * 16u * c.stride;                   // decode wavefronts' rows
    c.a_base = (uint32_t)(lane & 31) * c.stride + (uint32_t)hi * 16u;  // A operand of row tile 0
    c.ob = a.pcb;
    c.meta = __builtin_amdgcn_make_buffer_rsrc((void *)a.pslices, (short)0, (int)(a.nslices * 8),
                                               0x00020000 /* 32-bit data format */);
    if (wave >= mw) pk_dmain<E, MODE != 3>(c, wave - mw, s0);
    else if constexpr (MODE == 0) pk_mmain<KB>(c, qb, wave, s0);
    else if constexpr (MODE == 1) pk_mmainh<KB, 1>(c, qb, wave, s0);
    else if constexpr (MODE == 3) pk_mmains<KB>(c, qb, wave, s0);
    else pk_mmainh<KB, 2>(c, qb, 2 * wave, s0);
}

#define PK_CASES(KB, MODE)                       \
    switch (E) {                                \
        case 18: return pk_scores<18, KB, MODE>; \
        case 19: return pk_scores<19, KB, MODE>; \
        case 20: return pk_scores<20, KB, MODE>; \
        case 21: return pk_scores<21, KB, MODE>; \
        case 22: return pk_scores<22, KB, MODE>; \
        case 23: return pk_scores<23, KB, MODE>; \
        default: return nullptr;                \
    }
