/*
 * gen.h — seeded synthetic operands (libsegtaco).
 *
 * No counterpart in the reference, whose inputs are MatrixMarket files with
 * time-seeded values (ds.h:270-274, timers.h:26-28) and B = ones
 * (ds.h:282-285). These generators produce the benchmark configurations
 * reproducibly: every draw is splitmix64 of (seed, stream, index).
 */
#ifndef SEGTACO_GEN_H
#define SEGTACO_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEGTACO_GEN_UNIFORM 0  /* round(mean) entries per row */
#define SEGTACO_GEN_POWERLAW 1 /* log-normal row lengths, capped */
#define SEGTACO_GEN_BANDED 2   /* |i - j| <= bw */

/* Writes pos[0..M] and returns nnz, or -1 on bad arguments / nnz > INT32_MAX.
 * mean, sigma, cap apply to POWERLAW (mean also to UNIFORM); bw to BANDED. */
int64_t segtaco_gen_pos(int32_t kind, uint64_t seed, int32_t M, int32_t K,
                        double mean, double sigma, int32_t cap, int32_t bw,
                        int32_t *pos);
/* Fills crd (unique, sorted per row) and vals (U[-1,1]); either may be NULL. */
int32_t segtaco_gen_fill(int32_t kind, uint64_t seed, int32_t M, int32_t K,
                         int32_t bw, const int32_t *pos, int32_t *crd,
                         float *vals);
/* out[i] = U[-1,1] for i < n. */
int32_t segtaco_gen_dense(uint64_t seed, int64_t n, float *out);
/* FP32 -> BF16 (dtype 1) or FP16 (dtype 2), round to nearest even. */
int32_t segtaco_convert_f32(int32_t dtype, int64_t n, const float *in,
                            uint16_t *out);

#ifdef __cplusplus
}
#endif
#endif
