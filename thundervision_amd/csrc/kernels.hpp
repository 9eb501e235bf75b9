// Launch helpers shared between the tdvh translation units.
#pragma once

#include "common.hpp"

namespace tdvh {

int launch_bt_terms(const uint8_t *L, const uint8_t *R, int W, int H, int pitch, uint32_t *TL,
                    uint32_t *TR, hipStream_t s);
int launch_bt_split(const uint32_t *TL, const uint32_t *TR, int W, int H, int D, uint32_t *out,
                    hipStream_t s);
int launch_convert(uint16_t *strided, int64_t sx, int64_t sy, uint32_t *split, int W, int H,
                   int D, bool to_split, hipStream_t s);

// BT term planes of the fused SGM path (splatted left terms, paired right
// terms; see sgm.hip).
int launch_bt_planes(const uint8_t *L, const uint8_t *R, int W, int H, int pitch, int m,
                     uint32_t *LT, uint32_t *RT, hipStream_t s);

// Reference-geometry SGM into an int32 strided [y][x][d] volume.
int run_sgm_ref(tdvh_ctx *ctx, const uint32_t *C_split, int W, int H, int D, int P1, int P2,
                int flags, int32_t *S, hipStream_t s);

// SGM over a strided uint32 or float volume into S (same dtype, strided),
// directions of p->dir_mask in increasing index (sgm_wide.hip).
int run_sgm_wide(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *p, const tdvh_dsi *S,
                 hipStream_t s);

int popcount8(int mask);

// Frees the reference-geometry tables cached for a context.
void release_ref_tables(const tdvh_ctx *ctx);

} // namespace tdvh
