/*
 * Seeded synthetic stereo-pair generator (workload inputs for tests and the
 * benchmark; the reference ships no generator).
 *
 * Recipe:
 *   left   uniform uint8 noise -> separable Gaussian blur (sigma 1.5,
 *          radius 5, reflect-101 borders, double precision) -> contrast
 *          stretch to [0, 255], rounded to nearest;
 *   truth  background disparity plus nrect axis-aligned rectangles, each
 *          with an integer disparity uniform in [0, dmax]; later rectangles
 *          overwrite earlier ones; with frame > 0 each rectangle drifts by
 *          its own velocity of 1..3 px per frame on each axis (bouncing at
 *          the image border);
 *   right  R(x - d, y) = L(x, y) forward-warped with the larger disparity
 *          winning; right pixels no left pixel lands on (occluded or out of
 *          range) get uniform noise; then additive Gaussian noise
 *          (sigma 2, Box-Muller), rounded and clipped to uint8.
 * The PRNG is SplitMix64-seeded xoshiro256**, so both hosts produce the same
 * bytes for the same (seed, frame).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} tdv_rng;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(tdv_rng *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(tdv_rng *r)
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static double rng_unit(tdv_rng *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

/* uniform integer in [lo, hi] */
static int rng_int(tdv_rng *r, int lo, int hi)
{
    if (hi <= lo) return lo;
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    return lo + (int)(rng_next(r) % span);
}

static double rng_gauss(tdv_rng *r)
{
    const double u1 = 1.0 - rng_unit(r); /* (0, 1] */
    const double u2 = rng_unit(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static int reflect101(int i, int n)
{
    if (n == 1) return 0;
    for (;;) {
        if (i < 0)
            i = -i;
        else if (i >= n)
            i = 2 * n - 2 - i;
        else
            return i;
    }
}

/* position of a drifting rectangle edge: bounce inside [0, span] */
static int bounce(int p0, int v, int frame, int span)
{
    if (span <= 0) return 0;
    const long long period = 2LL * span;
    long long p = ((long long)p0 + (long long)v * frame) % period;
    if (p < 0) p += period;
    return (int)(p <= span ? p : period - p);
}

int tdv_synth_pair(int W, int H, int dmax, int nrect, uint64_t seed, int frame,
                   uint8_t *L, uint8_t *R, int16_t *gt)
{
    if (W <= 0 || H <= 0 || dmax < 0 || nrect < 0) return -1;
    const size_t n = (size_t)W * H;
    tdv_rng rng;
    rng_seed(&rng, seed);

    /* left texture */
    double *a = (double *)malloc(sizeof(double) * n);
    double *b = (double *)malloc(sizeof(double) * n);
    int16_t *truth = gt ? gt : (int16_t *)malloc(sizeof(int16_t) * n);
    if (!a || !b || !truth) {
        free(a);
        free(b);
        if (!gt) free(truth);
        return -2;
    }
    for (size_t i = 0; i < n; i++) a[i] = (double)(rng_next(&rng) >> 56);
    const double sigma = 1.5;
    double k[11], ksum = 0.0;
    for (int i = -5; i <= 5; i++) {
        k[i + 5] = exp(-(double)(i * i) / (2.0 * sigma * sigma));
        ksum += k[i + 5];
    }
    for (int i = 0; i < 11; i++) k[i] /= ksum;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double s = 0.0;
            for (int i = -5; i <= 5; i++) s += k[i + 5] * a[(size_t)y * W + reflect101(x + i, W)];
            b[(size_t)y * W + x] = s;
        }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double s = 0.0;
            for (int i = -5; i <= 5; i++) s += k[i + 5] * b[(size_t)reflect101(y + i, H) * W + x];
            a[(size_t)y * W + x] = s;
        }
    double mn = a[0], mx = a[0];
    for (size_t i = 1; i < n; i++) {
        if (a[i] < mn) mn = a[i];
        if (a[i] > mx) mx = a[i];
    }
    for (size_t i = 0; i < n; i++) {
        const double v = (mx > mn) ? (a[i] - mn) / (mx - mn) * 255.0 : 128.0;
        int q = (int)floor(v + 0.5);
        L[i] = (uint8_t)(q < 0 ? 0 : (q > 255 ? 255 : q));
    }

    /* ground-truth disparity */
    const int bg = rng_int(&rng, 0, dmax / 2);
    for (size_t i = 0; i < n; i++) truth[i] = (int16_t)bg;
    for (int r = 0; r < nrect; r++) {
        const int wmin = W / 16 + 1, wmax = W / 4 > wmin ? W / 4 : wmin;
        const int hmin = H / 16 + 1, hmax = H / 4 > hmin ? H / 4 : hmin;
        int w = rng_int(&rng, wmin, wmax), h = rng_int(&rng, hmin, hmax);
        if (w > W) w = W;
        if (h > H) h = H;
        const int x0 = rng_int(&rng, 0, W - w), y0 = rng_int(&rng, 0, H - h);
        const int d = rng_int(&rng, 0, dmax);
        const int vx = rng_int(&rng, 1, 3) * ((rng_next(&rng) & 1) ? 1 : -1);
        const int vy = rng_int(&rng, 1, 3) * ((rng_next(&rng) & 1) ? 1 : -1);
        const int xs = bounce(x0, vx, frame, W - w), ys = bounce(y0, vy, frame, H - h);
        for (int y = ys; y < ys + h; y++)
            for (int x = xs; x < xs + w; x++) truth[(size_t)y * W + x] = (int16_t)d;
    }

    /* right image: forward warp with z-buffer, per-frame noise stream */
    tdv_rng rn;
    rng_seed(&rn, seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(frame + 1)));
    int16_t *zb = (int16_t *)a; /* reuse: n int16 fit in n doubles */
    for (size_t i = 0; i < n; i++) zb[i] = -1;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int d = truth[(size_t)y * W + x];
            const int xr = x - d;
            if (xr < 0) continue;
            const size_t o = (size_t)y * W + xr;
            if (d > zb[o]) {
                zb[o] = (int16_t)d;
                R[o] = L[(size_t)y * W + x];
            }
        }
    for (size_t i = 0; i < n; i++)
        if (zb[i] < 0) R[i] = (uint8_t)(rng_next(&rn) >> 56);
    for (size_t i = 0; i < n; i++) {
        const double v = (double)R[i] + 2.0 * rng_gauss(&rn);
        int q = (int)floor(v + 0.5);
        R[i] = (uint8_t)(q < 0 ? 0 : (q > 255 ? 255 : q));
    }
    free(a);
    free(b);
    if (!gt) free(truth);
    return 0;
}

/* Uniform integer in [lo, hi] from the stream (seed, stream); used to draw
 * per-pair shapes of mixed batches reproducibly. */
int tdv_synth_uniform_int(uint64_t seed, uint64_t stream, int lo, int hi)
{
    tdv_rng r;
    rng_seed(&r, seed * 0x9E3779B97F4A7C15ull + stream);
    return rng_int(&r, lo, hi);
}

/* Exponential inter-arrival time (seconds) for Poisson arrivals at `rate`. */
double tdv_synth_exponential(uint64_t seed, uint64_t stream, double rate)
{
    tdv_rng r;
    rng_seed(&r, seed * 0x9E3779B97F4A7C15ull + stream);
    return -log(1.0 - rng_unit(&r)) / rate;
}

/* 64-bit FNV-1a over a buffer: used for generator known-answer hashes. */
uint64_t tdv_synth_hash(const uint8_t *p, size_t n)
{
    uint64_t h = 0xCBF29CE484222325ull;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}
