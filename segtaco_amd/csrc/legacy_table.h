// legacy_table.h — the segTACO operator surface (gpu_kernels.cuh:11-93) as
// (entry point, schedule, vec cap, col_lanes, group, tile) presets.
//
// vec is a cap (0 = widest the operands allow); zero fields are resolved from
// the operands. The reference's tuned kernels encode their parameters in the
// name; the mapping keeps what each parameter means on this hardware:
//   spmm_rb_pr_256_<c>_<g>_<gs> (spmm_rb_pr.h:428-948): g lanes split each
//     row's entries -> group = g slots; c consecutive columns per lane ->
//     vec cap c; gs (the size of the reference's atomic reduction group) has
//     no counterpart because the P slots are reduced without atomics.
//   spmm_eb_pr_<Nnz>_<gs> (spmm_eb_pr.h:482-779): Nnz entries per tile ->
//     tile = Nnz; gs lanes per segmented reduction -> group = gs slots.
//   spmm_rb_sr_256_<r>_<c> (spmm_rb_sr.h:351-535): r rows per thread ->
//     group = r rows per lane group; c columns per thread -> vec cap c.
//   spmm_eb_sr_256_<g>_<c> (spmm_eb_sr.h:856-1095): g entries per group of a
//     64*g*c-entry block -> tile = 64*g entries per lane group; c columns per
//     thread -> vec cap c.
// Unlike the reference's tuned kernels (which compute 4 or 128 columns only,
// spmm_rb_sr.h:363-376, spmm_csr_gpu_taco.h:133-134) every preset is correct
// for any N.
//
// The two vendor-library rows (spmm_csr_gpu_cusparse.h:52-260) run
// rocSPARSE's CSR SpMM instead of an engine schedule (SEGTACO_VENDOR_ROW):
// a comparison row for sweeps, not an engine dispatch target.
//
// X(name, schedule, vec_cap, col_lanes, group, tile)
#define SEGTACO_VENDOR_ROW 100
#define SEGTACO_LEGACY_TABLE(X)                                   \
  X(spmm_rb_sr, SEGTACO_RB_SR, 0, 0, 0, 0)                        \
  X(spmm_rb_pr, SEGTACO_RB_PR, 0, 0, 0, 0)                        \
  X(spmm_eb_sr, SEGTACO_EB_SR, 0, 0, 0, 0)                        \
  X(spmm_eb_pr, SEGTACO_EB_PR, 0, 0, 0, 0)                        \
  X(spmm_csr_gpu_taco, SEGTACO_EB_SR, 0, 0, 0, 256)               \
  X(spmm_csr_gpu_warp, SEGTACO_RB_SR, 4, 0, 0, 0)                 \
  X(spmm_csr_gpu_reduction, SEGTACO_RB_PR, 1, 1, 32, 0)           \
  X(spmm_csr_pr_eb, SEGTACO_EB_PR, 1, 1, 64, 256)                 \
  X(spmm_rb_sr_taco, SEGTACO_RB_SR, 4, 0, 0, 0)                   \
  X(spmm_rb_pr_taco, SEGTACO_RB_PR, 1, 1, 32, 0)                  \
  X(spmm_eb_sr_taco, SEGTACO_EB_SR, 0, 0, 0, 256)                 \
  X(spmm_eb_pr_taco, SEGTACO_EB_PR, 1, 1, 64, 256)                \
  X(spmm_eb_sr_search, SEGTACO_EB_SR, 0, 0, 0, 32)                \
  X(spmm_eb_pr_search, SEGTACO_EB_PR, 4, 0, 0, 32)                \
  X(spmm_rb_sr_512_64, SEGTACO_RB_SR, 2, 0, 4, 0)                 \
  X(spmm_eb_sr_512_64, SEGTACO_EB_SR, 2, 0, 0, 256)               \
  X(spmm_rb_pr_256, SEGTACO_RB_PR, 1, 1, 32, 0)                   \
  X(spmm_rb_pr_256_1_4_4, SEGTACO_RB_PR, 1, 0, 4, 0)              \
  X(spmm_rb_pr_256_2_4_4, SEGTACO_RB_PR, 2, 0, 4, 0)              \
  X(spmm_rb_pr_256_4_4_4, SEGTACO_RB_PR, 4, 0, 4, 0)              \
  X(spmm_rb_pr_256_1_8_4, SEGTACO_RB_PR, 1, 0, 8, 0)              \
  X(spmm_rb_pr_256_2_8_4, SEGTACO_RB_PR, 2, 0, 8, 0)              \
  X(spmm_rb_pr_256_4_8_4, SEGTACO_RB_PR, 4, 0, 8, 0)              \
  X(spmm_rb_pr_256_1_8_8, SEGTACO_RB_PR, 1, 0, 8, 0)              \
  X(spmm_rb_pr_256_2_8_8, SEGTACO_RB_PR, 2, 0, 8, 0)              \
  X(spmm_rb_pr_256_4_8_8, SEGTACO_RB_PR, 4, 0, 8, 0)              \
  X(spmm_rb_pr_256_1_16_4, SEGTACO_RB_PR, 1, 0, 16, 0)            \
  X(spmm_rb_pr_256_2_16_4, SEGTACO_RB_PR, 2, 0, 16, 0)            \
  X(spmm_rb_pr_256_4_16_4, SEGTACO_RB_PR, 4, 0, 16, 0)            \
  X(spmm_rb_pr_256_1_16_8, SEGTACO_RB_PR, 1, 0, 16, 0)            \
  X(spmm_rb_pr_256_2_16_8, SEGTACO_RB_PR, 2, 0, 16, 0)            \
  X(spmm_rb_pr_256_4_16_8, SEGTACO_RB_PR, 4, 0, 16, 0)            \
  X(spmm_rb_pr_256_1_16_16, SEGTACO_RB_PR, 1, 0, 16, 0)           \
  X(spmm_rb_pr_256_2_16_16, SEGTACO_RB_PR, 2, 0, 16, 0)           \
  X(spmm_rb_pr_256_4_16_16, SEGTACO_RB_PR, 4, 0, 16, 0)           \
  X(spmm_rb_pr_256_1_32_4, SEGTACO_RB_PR, 1, 0, 32, 0)            \
  X(spmm_rb_pr_256_2_32_4, SEGTACO_RB_PR, 2, 0, 32, 0)            \
  X(spmm_rb_pr_256_4_32_4, SEGTACO_RB_PR, 4, 0, 32, 0)            \
  X(spmm_rb_pr_256_1_32_8, SEGTACO_RB_PR, 1, 0, 32, 0)            \
  X(spmm_rb_pr_256_2_32_8, SEGTACO_RB_PR, 2, 0, 32, 0)            \
  X(spmm_rb_pr_256_4_32_8, SEGTACO_RB_PR, 4, 0, 32, 0)            \
  X(spmm_rb_pr_256_1_32_16, SEGTACO_RB_PR, 1, 0, 32, 0)           \
  X(spmm_rb_pr_256_2_32_16, SEGTACO_RB_PR, 2, 0, 32, 0)           \
  X(spmm_rb_pr_256_4_32_16, SEGTACO_RB_PR, 4, 0, 32, 0)           \
  X(spmm_rb_pr_256_1_32_32, SEGTACO_RB_PR, 1, 0, 32, 0)           \
  X(spmm_rb_pr_256_2_32_32, SEGTACO_RB_PR, 2, 0, 32, 0)           \
  X(spmm_rb_pr_256_4_32_32, SEGTACO_RB_PR, 4, 0, 32, 0)           \
  X(spmm_eb_pr_256_32, SEGTACO_EB_PR, 0, 0, 32, 256)              \
  X(spmm_eb_pr_256_16, SEGTACO_EB_PR, 0, 0, 16, 256)              \
  X(spmm_eb_pr_256_8, SEGTACO_EB_PR, 0, 0, 8, 256)                \
  X(spmm_eb_pr_256_4, SEGTACO_EB_PR, 0, 0, 4, 256)                \
  X(spmm_eb_pr_128_32, SEGTACO_EB_PR, 0, 0, 32, 128)              \
  X(spmm_eb_pr_128_16, SEGTACO_EB_PR, 0, 0, 16, 128)              \
  X(spmm_eb_pr_128_8, SEGTACO_EB_PR, 0, 0, 8, 128)                \
  X(spmm_eb_pr_128_4, SEGTACO_EB_PR, 0, 0, 4, 128)                \
  X(spmm_eb_pr_64_32, SEGTACO_EB_PR, 0, 0, 32, 64)                \
  X(spmm_eb_pr_64_16, SEGTACO_EB_PR, 0, 0, 16, 64)                \
  X(spmm_eb_pr_64_8, SEGTACO_EB_PR, 0, 0, 8, 64)                  \
  X(spmm_eb_pr_64_4, SEGTACO_EB_PR, 0, 0, 4, 64)                  \
  X(spmm_rb_sr_256_1_1, SEGTACO_RB_SR, 1, 0, 1, 0)                \
  X(spmm_rb_sr_256_1_2, SEGTACO_RB_SR, 2, 0, 1, 0)                \
  X(spmm_rb_sr_256_1_4, SEGTACO_RB_SR, 4, 0, 1, 0)                \
  X(spmm_rb_sr_256_2_1, SEGTACO_RB_SR, 1, 0, 2, 0)                \
  X(spmm_rb_sr_256_2_2, SEGTACO_RB_SR, 2, 0, 2, 0)                \
  X(spmm_rb_sr_256_2_4, SEGTACO_RB_SR, 4, 0, 2, 0)                \
  X(spmm_rb_sr_256_4_1, SEGTACO_RB_SR, 1, 0, 4, 0)                \
  X(spmm_rb_sr_256_4_2, SEGTACO_RB_SR, 2, 0, 4, 0)                \
  X(spmm_rb_sr_256_4_4, SEGTACO_RB_SR, 4, 0, 4, 0)                \
  X(spmm_eb_sr_256_4_1, SEGTACO_EB_SR, 1, 0, 0, 256)              \
  X(spmm_eb_sr_256_4_2, SEGTACO_EB_SR, 2, 0, 0, 256)              \
  X(spmm_eb_sr_256_4_4, SEGTACO_EB_SR, 4, 0, 0, 256)              \
  X(spmm_eb_sr_256_8_1, SEGTACO_EB_SR, 1, 0, 0, 512)              \
  X(spmm_eb_sr_256_8_2, SEGTACO_EB_SR, 2, 0, 0, 512)              \
  X(spmm_eb_sr_256_8_4, SEGTACO_EB_SR, 4, 0, 0, 512)              \
  X(spmm_eb_sr_256_16_1, SEGTACO_EB_SR, 1, 0, 0, 1024)            \
  X(spmm_eb_sr_256_16_2, SEGTACO_EB_SR, 2, 0, 0, 1024)            \
  X(spmm_eb_sr_256_16_4, SEGTACO_EB_SR, 4, 0, 0, 1024)            \
  X(spmm_csr_cusparse_row, SEGTACO_VENDOR_ROW, 0, 0, 0, 0)        \
  X(spmm_csr_gpu_cusparse, SEGTACO_VENDOR_ROW, 0, 0, 0, 0)

enum segtaco_legacy_id {
#define SEGTACO_LEGACY_ENUM(name, s, v, q, g, t) SEGTACO_LEGACY_##name,
  SEGTACO_LEGACY_TABLE(SEGTACO_LEGACY_ENUM)
#undef SEGTACO_LEGACY_ENUM
  SEGTACO_LEGACY_COUNT
};
