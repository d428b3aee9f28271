// sddmm.cpp — host side of the SDDMM entry point (include/segtaco/sddmm.h):
// operand checks, vector width / lane count selection, the shared row
// partition pass, and the kernel launch.
#include "engine.h"
#include "launch.h"
#include "sddmm.h"
#include "segtaco/sddmm.h"

namespace segtaco {
namespace {

int64_t sddmm_tiles(const segtaco_csr_t *S) {
  return S->nnz > 0 ? ceil_div(S->nnz, kSdTile) : 0;
}

bool aligned(const void *p, int64_t bytes) {
  return (reinterpret_cast<uintptr_t>(p) % (uintptr_t)bytes) == 0;
}

}  // namespace
}  // namespace segtaco

using namespace segtaco;

extern "C" {

size_t segtaco_sddmm_workspace_bytes(const segtaco_csr_t *S) {
  if (!S || S->nnz <= 0) return 0;
  return align_up(sizeof(int32_t) * (size_t)(sddmm_tiles(S) + 1), 256);
}

int32_t segtaco_sddmm(const segtaco_csr_t *S, const void *X, int64_t ldx,
                      const void *Y, int64_t ldy, int32_t J, int32_t dtype,
                      float *out, void *workspace, size_t workspace_bytes,
                      void *stream_v) {
  if (!S || S->M < 0 || S->K < 0 || S->nnz < 0 || J < 0)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_sddmm: bad operand shape");
  if (S->nnz > kMaxNnz)
    return set_error(SEGTACO_ERR_UNSUPPORTED, "segtaco_sddmm: nnz %d above the supported %d",
                     S->nnz, kMaxNnz);
  if (dtype < SEGTACO_DT_F32 || dtype > SEGTACO_DT_F16)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_sddmm: unknown dtype %d", dtype);
  if (S->vals && S->vals_dtype != dtype)
    return set_error(SEGTACO_ERR_INVALID,
                     "segtaco_sddmm: values stored as %d but X/Y as %d", S->vals_dtype,
                     dtype);
  if (S->nnz == 0) return SEGTACO_OK;
  if (S->M == 0 || !S->pos || !S->crd || !out)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_sddmm: null operand");
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  if (J == 0) {
    SEGTACO_HIP(launch_fill_zero(out, 1, S->nnz, S->nnz, stream));
    return SEGTACO_OK;
  }
  if (!X || !Y || ldx < J || ldy < J)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_sddmm: X/Y missing or leading dimension < J");
  const size_t need = segtaco_sddmm_workspace_bytes(S);
  if (!workspace || workspace_bytes < need)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_sddmm: workspace of %zu bytes needed",
                     need);
  // widest vector that J, both leading dimensions and both bases allow
  const int32_t eb = elem_bytes(dtype);
  int32_t V = dtype == SEGTACO_DT_F32 ? 4 : 8;
  while (V > 1 && (J % V || ldx % V || ldy % V || !aligned(X, (int64_t)V * eb) ||
                   !aligned(Y, (int64_t)V * eb)))
    V >>= 1;
  int32_t lq = 0;
  while ((1 << lq) < 64 && (int64_t)(1 << lq) * V < J) ++lq;

  SddmmArgs a{};
  a.M = S->M;
  a.K = S->K;
  a.J = J;
  a.nnz = S->nnz;
  a.pos = S->pos;
  a.crd = S->crd;
  a.vals = S->vals;
  a.X = X;
  a.ldx = ldx;
  a.Y = Y;
  a.ldy = ldy;
  a.out = out;
  a.lq = lq;
  a.off32 = b_off32_ok(S->M, ldx, eb) && b_off32_ok(S->K, ldy, eb);
  a.tile = kSdTile;
  a.ntiles = (int32_t)sddmm_tiles(S);
  int32_t *tile_row = static_cast<int32_t *>(workspace);
  a.tile_row = tile_row;
  SEGTACO_HIP(launch_tile_rows(S->pos, S->M, kSdTile, a.ntiles, tile_row, stream));
  SEGTACO_HIP(launch_sddmm(dtype, V, a, stream));
  return SEGTACO_OK;
}

}  // extern "C"
