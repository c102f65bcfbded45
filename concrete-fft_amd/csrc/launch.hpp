// launch.hpp -- host-side launchers of the gfx950 kernels (C++ linkage,
// internal to libcfft_amd.so).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "program.hpp"

namespace cfft_launch {

struct dc64 {
  double re, im;
};

// Largest transform chunk held in one workgroup's LDS (points of 16 B).
constexpr uint32_t kLdsMaxLog2 = 13;

// Runs `prog` on `ntrans` chunks of 2^prog.log2n points; chunk t lives at
// data + (t >> chunks_log2) * stride + (t & (2^chunks_log2 - 1)) * 2^log2n.
hipError_t fft_lds(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans,
                   uint32_t chunks_log2, uint64_t stride, const dc64 *tw, hipStream_t s);

// One DEPTH_FWD / DEPTH_INV pass straight over global memory (levels larger
// than an LDS chunk).
hipError_t depth_global(const cfft_prog::Pass &P, dc64 *data, uint64_t batch, uint32_t log2n,
                        uint64_t stride, const dc64 *tw, hipStream_t s);

// One pass of any kind over global memory, src -> dst (dst may equal src for
// END and DEPTH_* passes, whose outputs land where their inputs were).
hipError_t pass_global(const cfft_prog::Pass &P, const dc64 *src, uint64_t src_stride, dc64 *dst,
                       uint64_t dst_stride, uint64_t batch, uint32_t log2n, const dc64 *tw,
                       hipStream_t s);

// buf[b*stride + i] = T[(freq[i] * degree) & (n-1)]
hipError_t monomial(dc64 *buf, uint32_t log2n, uint64_t degree, const uint32_t *freq,
                    const dc64 *tw, hipStream_t s);

// dst[b*n + i] = src[b*n + idx[i]]   /   dst[b*n + idx[i]] = src[b*n + i]
hipError_t gather(const dc64 *src, dc64 *dst, const uint32_t *idx, uint32_t n, uint64_t batch,
                  hipStream_t s);
hipError_t scatter(const dc64 *src, dc64 *dst, const uint32_t *idx, uint32_t n, uint64_t batch,
                   hipStream_t s);

// Compile-time pass programs (fixed_exec.hpp): launch the plan's product /
// whole-transform program on a kernel specialised to it and return true, or
// return false (nothing launched) when no such kernel is built for it.
// *e: the launch status.
bool fixed_polymul_n512(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                        const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                        const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                        hipError_t *e);
bool fixed_polymul_n1024(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                         const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                         const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                         hipError_t *e);
bool fixed_polymul_n2048(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                         const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                         const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                         hipError_t *e);
bool fixed_polymul_n4096(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                         const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                         const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                         hipError_t *e);
bool fixed_fft_n4096(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans,
                     uint64_t stride, const dc64 *tw, hipStream_t s, hipError_t *e);
bool fixed_polymul_n8192(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                         const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                         const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                         hipError_t *e);
bool fixed_fft_n8192(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans,
                     uint64_t stride, const dc64 *tw, hipStream_t s, hipError_t *e);
bool fixed_polymul_n256(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                        const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                        const dc64 *tw, const dc64 *twist, const dc64 *untwist, hipStream_t s,
                        hipError_t *e);
bool fixed_fft_n256(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans, uint64_t stride,
                    const dc64 *tw, hipStream_t s, hipError_t *e);
bool fixed_fft_n512(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans, uint64_t stride,
                    const dc64 *tw, hipStream_t s, hipError_t *e);
bool fixed_fft_n1024(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans,
                     uint64_t stride, const dc64 *tw, hipStream_t s, hipError_t *e);
bool fixed_fft_n2048(const cfft_prog::Program &prog, dc64 *data, uint64_t ntrans,
                     uint64_t stride, const dc64 *tw, hipStream_t s, hipError_t *e);

// Fused negacyclic product over the generic pass executor (a compile-time
// program's kernel where one is built).
hipError_t polymul_generic(const cfft_prog::Program &fwd, const cfft_prog::Program &inv,
                           const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                           const dc64 *tw, const dc64 *twist, const dc64 *untwist,
                           hipStream_t s);

// Register-resident kernels for Dif16 plans with base_n == n (r16_exec.hpp).
bool r16_supported(uint32_t log2n);
hipError_t r16_transform(uint32_t log2n, bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                         const dc64 *tw, hipStream_t s);
hipError_t r16_polymul(uint32_t log2n, const uint64_t *a, const int64_t *b, uint64_t *c,
                       uint64_t batch, const dc64 *tw, const dc64 *twist, hipStream_t s);


// Dif16 plans with n = base_n = 512: four transforms per wavefront (two per
// 32-lane half, r16l_transform<9>); tw_blk is the 1024-point r16 block.
hipError_t w512_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                          const dc64 *tw_blk, hipStream_t s);

// Negacyclic product (N = 512 / 1024) on a Dif16 plan with n = base_n = 256
// or 512 (log2n 8, 9): one product per group of n / 16 lanes; tw_blk is the
// plan's 1024-point r16 block.
hipError_t wsub_polymul(uint32_t log2n, const uint64_t *a, const int64_t *b, uint64_t *c,
                        uint64_t batch, const dc64 *tw_blk, const dc64 *twist, hipStream_t s);

// Dif16 plans with n = 1024 and base_n = 512: one transform per 32-lane
// half-wave (r16l_depth2_transform<9>); tw is the plan's table, tw_blk its
// 1024-point r16 block.
hipError_t wd2_512_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                             const dc64 *tw, const dc64 *tw_blk, hipStream_t s);
// the fused negacyclic product (N = 2048) on the same plans: one product per
// wavefront, the depth pass's partner values swapped between its half-waves
hipError_t wd2_512_polymul(const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                           const dc64 *tw, const dc64 *tw_blk, const dc64 *twist,
                           hipStream_t s);

// Dif16 plans with n = base_n = 1024: two transforms per wavefront
// (r16w_transform), 8 wavefronts per workgroup; tw_r16 is the plan's r16
// twiddle block.
hipError_t w1024_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                           const dc64 *tw_r16, hipStream_t s);

// Dif16 plans with n = 2048 and base_n = 1024: one transform per wavefront
// (r16l_2048_transform); tw is the plan's table, tw_blk its 1024-point block.
// Dif16 plans with base_n = 1024 and n = 4096 (log2n 12): the plan's depth
// pass in registers around one 1024-point base transform per wave, two
// transforms per workgroup (tw: the plan's n-point table, tw_blk: its
// 1024-point r16 block).
hipError_t depth_transform(bool fwd, uint32_t log2n, dc64 *data, uint64_t batch, uint64_t stride,
                           const dc64 *tw, const dc64 *tw_blk, hipStream_t s);
// Fused negacyclic product on Dif16 plans with base_n = 1024 and n = 2048,
// 4096 or 8192 (log2n 11 .. 13): the depth-pass structure above for both
// operands and the inverse (twist / untwist tables of the plan).
hipError_t depth_polymul(uint32_t log2n, const uint64_t *a, const int64_t *b, uint64_t *c,
                         uint64_t batch, const dc64 *tw, const dc64 *tw_blk, const dc64 *twist,
                         const dc64 *untwist, hipStream_t s);
hipError_t w2048_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride, const dc64 *tw,
                           const dc64 *tw_blk, hipStream_t s);
// Dif16 plans with n = 16384 and base_n = 1024 in one pass: both depth
// levels in registers, 16 base transforms per workgroup (tw: the plan's
// 16384-point table, tw2: its 2048-point sub-table, tw_blk: the 1024-point
// r16 block).
hipError_t depth16k_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                              const dc64 *tw, const dc64 *tw2, const dc64 *tw_blk,
                              hipStream_t s);
// Dif16 plans with n = 8192 and base_n = 1024: the plan's radix-8 depth level
// in registers, two transforms per workgroup through LDS in turn.
hipError_t depth8k_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                             const dc64 *tw, const dc64 *tw_blk, hipStream_t s);

// Register-resident kernels for Dit16 plans with base_n == n == 1024
// (r16dit_exec.hpp): transforms (in place, lane-linear order) and the fused
// negacyclic product at N = 2048; tw_r16 is the plan's r16 twiddle block.
hipError_t dit16_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride,
                           const dc64 *tw_r16, hipStream_t s);
hipError_t dit16_polymul(const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                         const dc64 *tw_r16, const dc64 *twist, hipStream_t s);

// Register-resident kernels for Dif8 plans with base_n == n == 1024 (r8_exec.hpp).
bool r8_supported(uint32_t log2n);
hipError_t r8_transform(bool fwd, dc64 *data, uint64_t batch, uint64_t stride, const dc64 *tw,
                        hipStream_t s);
hipError_t r8_polymul(const uint64_t *a, const int64_t *b, uint64_t *c, uint64_t batch,
                      const dc64 *tw, const dc64 *twist, hipStream_t s);

// Row copies between a contiguous block and caller buffers named by a table
// of device addresses (`table`: `rows` u64 addresses, itself readable by the
// device): gather copies row r from table[r] to dst + r * row_words, scatter
// from src + r * row_words to table[r].  rows <= 65535.
hipError_t gather_rows(const uint64_t *table, uint64_t rows, uint64_t row_words, uint64_t *dst,
                       hipStream_t s);
hipError_t scatter_rows(const uint64_t *table, uint64_t rows, uint64_t row_words,
                        const uint64_t *src, hipStream_t s);

}  // namespace cfft_launch
