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
// SGM over uint32 / float volumes: S into *S, or (S null, one-launch forms
// only) the lowest-index WTA's disparities straight from the sum.
bool sgm_wide_one_launch(const tdvh_dsi *C, const tdvh_sgm_params *p);
int run_sgm_wide(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *p, const tdvh_dsi *S,
                 int16_t *disp, int disp_pitch, hipStream_t s);

int popcount8(int mask);

// Whether the SGM scans of a W x H pair at m words per lane and penalty P2
// read relative cost bytes (sgm_rel_pixel) rather than recomputing BT
// (fused) or reading the uint16 volume (volume true): the choice of the
// per-pair path, sgm.hip.
bool sgm_rel_auto(int W, int H, int m, int p2, bool volume);

// Appends the line-group table of one W x H pair (entries plane slot << 28 |
// direction << 24 | first line; lpw lines per group) to `out`, ordered
// longest first within chiplet bands for wpg waves per workgroup, entry k of
// the launch running on workgroup k / wpg (sgm.hip).
void line_order_append(int W, int H, int dir_mask, int lpw, int wpg, std::vector<uint32_t> &out);

// Frees the reference-geometry tables cached for a context.
void release_ref_tables(const tdvh_ctx *ctx);

} // namespace tdvh
