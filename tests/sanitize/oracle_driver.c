/* Runs every function of the CPU oracle on small and degenerate inputs
 * (1x1 images, D = 1, D above the width, tall and wide images) so that a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer checks its
 * indexing and arithmetic. Prints "oracle sanitize ok". */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tdvref.h"

static uint64_t rng = 88172645463325252ull;
static uint32_t next_u32(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)(rng >> 16);
}

static void run_case(int W, int H, int D)
{
    const size_t px = (size_t)W * H, n = px * D;
    uint8_t *L = malloc(px), *R = malloc(px), *u8 = malloc(px), *rgb = malloc(px * 4);
    uint16_t *C16 = malloc(n * 2), *S16 = malloc(n * 2), *Lr = malloc(n * 2);
    uint32_t *C32 = malloc(n * 4), *S32 = malloc(n * 4);
    int32_t *Si = malloc(n * 4), *vis = malloc(px * 8 * 4);
    double *Cf64 = malloc(n * 8), *rho = malloc(n * 8);
    float *Cf = malloc(n * 4), *Sf = malloc(n * 4), *mx = malloc(px * 4), *my = malloc(px * 4);
    float *xyz = malloc(px * 12);
    int16_t *disp = malloc(px * 2), *disp2 = malloc(px * 2);
    uint64_t *acc = malloc((size_t)H * D * 8);
    for (size_t i = 0; i < px; i++) {
        L[i] = (uint8_t)next_u32();
        R[i] = (uint8_t)next_u32();
        mx[i] = (float)(next_u32() % (4 * W + 4)) * 0.37f - 1.0f;
        my[i] = (float)(next_u32() % (4 * H + 4)) * 0.29f - 1.0f;
    }
    for (size_t i = 0; i < px * 4; i++) rgb[i] = (uint8_t)next_u32();
    tdvref_cost_bt2(L, R, W, H, D, C16);
    tdvref_cost_ssd(L, R, W, H, D, 1, C32);
    tdvref_cost_ssd_box(L, R, W, H, D, 2, C32);
    tdvref_cost_ncc(L, R, W, H, D, 1, 1, Cf64, rho);
    tdvref_cost_ncc(L, R, W, H, D, 2, 0, Cf64, NULL);
    for (int dir = 0; dir < 8; dir++) tdvref_sgm_lr(C16, W, H, D, 8, 32, dir, Lr);
    tdvref_sgm_std(C16, W, H, D, 8, 32, 0xFF, S16);
    tdvref_sgm_u32(C32, W, H, D, 8, 32, 0xFF, S32);
    for (size_t i = 0; i < n; i++) Cf[i] = (float)Cf64[i];
    tdvref_sgm_f32(Cf, W, H, D, 0.1f, 0.5f, 0xFF, Sf);
    const size_t np = tdvref_sgm_path_count(W, H);
    tdvref_sgpath *paths = malloc(np * sizeof(tdvref_sgpath));
    if (tdvref_sgm_paths(W, H, paths) != np) abort();
    free(paths);
    tdvref_sgm_ref(C16, W, H, D, 8, 32, TDVREF_SGM_START_ASSIGN | TDVREF_SGM_MIN_SKIP_LAST, Si);
    tdvref_sgm_ref_visits(W, H, vis);
    tdvref_wta_u16(S16, W, H, D, disp);
    tdvref_wta_u32(S32, W, H, D, disp);
    tdvref_wta_i32(Si, W, H, D, disp);
    tdvref_wta_f64(Cf64, W, H, D, disp);
    tdvref_wta_f32(Sf, W, H, D, disp);
    if ((D & (D - 1)) == 0) tdvref_wta_tree_u16(S16, W, H, D, disp);
    tdvref_dp(C32, W, H, D, 20, 0, disp, acc);
    tdvref_dp(C32, W, H, D, 20, TDVREF_DP_START_ZERO | TDVREF_DP_REF_BACKTRACK, disp, NULL);
    tdvref_dp_f32(Cf, W, H, D, 0.25f, 0, disp);
    tdvref_fast_wta_ssd(L, R, W, H, D, 3, disp);
    tdvref_gauss5_u8(L, W, H, u8);
    tdvref_open_i16(disp, W, H, 3, disp2);
    tdvref_rgb_to_gray_u8(rgb, W, H, 3, 0, u8);
    tdvref_rgb_to_gray_u8(rgb, W, H, 4, 1, u8);
    tdvref_remap_u8(L, W, H, mx, my, W, H, 7, u8);
    const float Q[16] = {1, 0, 0, -W / 2.0f, 0, 1, 0, -H / 2.0f, 0, 0, 0, 700, 0, 0, 8, 0};
    tdvref_reproject(disp, W, H, Q, 1.0f / D, 24.0f, xyz);
    free(L), free(R), free(u8), free(rgb), free(C16), free(S16), free(Lr), free(C32);
    free(S32), free(Si), free(vis), free(Cf64), free(rho), free(Cf), free(Sf), free(mx);
    free(my), free(xyz), free(disp), free(disp2), free(acc);
}

int main(void)
{
    const int shapes[][3] = {{1, 1, 1}, {1, 7, 3}, {9, 1, 4}, {5, 4, 8},  {13, 11, 16},
                             {3, 17, 6}, {31, 9, 33}, {8, 8, 1}, {2, 2, 5}};
    for (size_t i = 0; i < sizeof shapes / sizeof shapes[0]; i++)
        run_case(shapes[i][0], shapes[i][1], shapes[i][2]);
    printf("oracle sanitize ok\n");
    return 0;
}
