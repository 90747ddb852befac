/* Seeded synthetic RGB-D pair generator.  Interface and scene model: pdflow_synth.h.
 *
 * Randomness: SplitMix64 (Steele, Lea, Flood 2014; the 64-bit finaliser of
 * Vigna's reference description).  Scene parameters come from one sequential
 * SplitMix64 stream seeded by the pair seed; per-pixel noise and holes come from a
 * counter hash of (seed, frame, pixel), so the output does not depend on the order
 * in which pixels are visited.
 */
#include "pdflow_synth.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAX_PATCHES 12
#define NUM_WAVES 7

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct { uint64_t state; } rng_t;

static uint64_t rng_next(rng_t *r)
{
    r->state += 0x9E3779B97F4A7C15ull;
    return mix64(r->state);
}

/* uniform double in [0, 1) with 53 random bits */
static double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static double rng_range(rng_t *r, double lo, double hi) { return lo + (hi - lo) * rng_unit(r); }

/* counter-based uniform in [0,1) for (seed, stream, index) */
static double hash_unit(uint64_t seed, uint64_t stream, uint64_t index)
{
    uint64_t z = mix64(seed ^ 0xD1B54A32D192ED03ull);
    z = mix64(z ^ (stream * 0x9E3779B97F4A7C15ull));
    z = mix64(z ^ (index + 0x632BE59BD9B4E019ull));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct { double x, y, z; } vec3;

static vec3 v3(double x, double y, double z) { vec3 r = {x, y, z}; return r; }
static vec3 vadd(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static vec3 vsub(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static vec3 vscale(vec3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static double vdot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static vec3 vcross(vec3 a, vec3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static vec3 vnorm(vec3 a) { return vscale(a, 1.0 / sqrt(vdot(a, a))); }

/* rotation of a by angle th about unit axis k (Rodrigues) */
static vec3 vrot(vec3 a, vec3 k, double th)
{
    const double c = cos(th), s = sin(th);
    return vadd(vadd(vscale(a, c), vscale(vcross(k, a), s)), vscale(k, vdot(k, a) * (1.0 - c)));
}

typedef struct {
    vec3 p0, n, e1, e2;       /* centre, normal (towards the camera), in-plane axes */
    double half1, half2;      /* half extents along e1/e2; <= 0 means unbounded */
    double base, amp[NUM_WAVES], fs[NUM_WAVES], ft[NUM_WAVES], ph[NUM_WAVES];
} patch_t;

typedef struct {
    vec3 axis, trans;
    double angle;
} motion_t;

static void make_scene(uint64_t seed, unsigned rows, unsigned cols, patch_t *patches,
                       motion_t *mot, int *npatch)
{
    rng_t r = {seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
    const double f = 525.0 * cols / 640.0;
    const double cx = 0.5 * (cols - 1.0), cy = 0.5 * (rows - 1.0);
    const int n = 6 + (int)(rng_next(&r) % 7u); /* 6..12 patches */
    *npatch = n;

    for (int i = 0; i < n; i++) {
        patch_t *p = &patches[i];
        double depth, tilt_x, tilt_y;
        double pu, pv;
        if (i == 0) { /* background: unbounded, far */
            depth = rng_range(&r, 3.2, 3.9);
            tilt_x = rng_range(&r, -0.35, 0.35);
            tilt_y = rng_range(&r, -0.25, 0.25);
            pu = cx; pv = cy;
        } else {
            depth = rng_range(&r, 0.8, 3.0);
            tilt_x = rng_range(&r, -0.7, 0.7);
            tilt_y = rng_range(&r, -0.7, 0.7);
            pu = rng_range(&r, 0.1 * cols, 0.9 * cols);
            pv = rng_range(&r, 0.1 * rows, 0.9 * rows);
        }
        p->p0 = v3((pu - cx) / f * depth, (pv - cy) / f * depth, depth);
        /* normal: towards the camera (-z), tilted about x and y */
        p->n = vnorm(v3(sin(tilt_x), sin(tilt_y), -1.0));
        vec3 up = v3(0.0, 1.0, 0.0);
        p->e1 = vnorm(vcross(up, p->n));
        p->e2 = vcross(p->n, p->e1);
        if (i == 0) {
            p->half1 = p->half2 = -1.0;
        } else {
            p->half1 = rng_range(&r, 0.12, 0.45) * depth;
            p->half2 = rng_range(&r, 0.12, 0.45) * depth;
        }
        p->base = rng_range(&r, 70.0, 185.0);
        double amp_total = rng_range(&r, 45.0, 75.0);
        for (int w = 0; w < NUM_WAVES; w++) {
            const double fmag = rng_range(&r, 1.5, 14.0);      /* cycles per metre */
            const double dir = rng_range(&r, 0.0, 2.0 * M_PI);
            p->fs[w] = fmag * cos(dir);
            p->ft[w] = fmag * sin(dir);
            p->ph[w] = rng_range(&r, 0.0, 2.0 * M_PI);
            p->amp[w] = amp_total / NUM_WAVES * rng_range(&r, 0.5, 1.5);
        }
        /* rigid motion: rotation <= 2 deg about a random axis through p0, |t| <= 5 cm */
        vec3 axis;
        do {
            axis = v3(rng_range(&r, -1, 1), rng_range(&r, -1, 1), rng_range(&r, -1, 1));
        } while (vdot(axis, axis) < 1e-4 || vdot(axis, axis) > 1.0);
        mot[i].axis = vnorm(axis);
        mot[i].angle = rng_range(&r, 0.0, 2.0) * M_PI / 180.0;
        vec3 t;
        do {
            t = v3(rng_range(&r, -1, 1), rng_range(&r, -1, 1), rng_range(&r, -1, 1));
        } while (vdot(t, t) > 1.0);
        mot[i].trans = vscale(t, 0.05);
    }
}

static patch_t moved_patch_t(const patch_t *p, const motion_t *m, double t)
{
    /* the motion applied t times: rotation t*angle about the same axis, translation t*trans */
    patch_t q = *p;
    q.p0 = vadd(p->p0, vscale(m->trans, t));
    q.n = vrot(p->n, m->axis, m->angle * t);
    q.e1 = vrot(p->e1, m->axis, m->angle * t);
    q.e2 = vrot(p->e2, m->axis, m->angle * t);
    return q;
}

static patch_t moved_patch(const patch_t *p, const motion_t *m) { return moved_patch_t(p, m, 1.0); }

/* ray-casts the patches for one pixel; returns the hit patch or -1, depth in *z,
 * local coordinates in *s, *t */
static int cast(const patch_t *ps, int n, vec3 d, double *z, double *s, double *t)
{
    int best = -1;
    double bz = 1e30, bs = 0, bt = 0;
    for (int i = 0; i < n; i++) {
        const patch_t *p = &ps[i];
        const double den = vdot(p->n, d);
        if (fabs(den) < 1e-9) continue;
        const double lam = vdot(p->n, p->p0) / den;
        if (lam <= 0.0 || lam >= bz) continue;
        vec3 x = vscale(d, lam);
        vec3 rel = vsub(x, p->p0);
        const double ls = vdot(rel, p->e1), lt = vdot(rel, p->e2);
        if (p->half1 > 0.0 && (fabs(ls) > p->half1 || fabs(lt) > p->half2)) continue;
        best = i; bz = lam; bs = ls; bt = lt;
    }
    *z = bz; *s = bs; *t = bt;
    return best;
}

static double texture(const patch_t *p, double s, double t)
{
    double val = p->base;
    for (int w = 0; w < NUM_WAVES; w++)
        val += p->amp[w] * sin(2.0 * M_PI * (p->fs[w] * s + p->ft[w] * t) + p->ph[w]);
    return val;
}

static uint8_t to_u8(double x)
{
    double r = floor(x + 0.5);
    if (r < 0.0) r = 0.0;
    if (r > 255.0) r = 255.0;
    return (uint8_t)r;
}

static uint16_t to_u16_depth(double z)
{
    if (!(z >= 0.5 && z <= 4.5)) return 0;
    return (uint16_t)floor(z * 5000.0 + 0.5);
}

/* standard normal from two counter-hash uniforms (Box-Muller) */
static double gauss(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u1 = hash_unit(seed, stream, 2 * idx);
    const double u2 = hash_unit(seed, stream, 2 * idx + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static void render(uint64_t seed, unsigned frame, const patch_t *ps, int n, unsigned rows,
                   unsigned cols, int noise, uint8_t *I, uint16_t *Z, int *hit, double *zout)
{
    const double f = 525.0 * cols / 640.0;
    const double cx = 0.5 * (cols - 1.0), cy = 0.5 * (rows - 1.0);
    for (unsigned u = 0; u < cols; u++)
        for (unsigned v = 0; v < rows; v++) {
            const unsigned idx = v + u * rows;
            vec3 d = v3((u - cx) / f, (v - cy) / f, 1.0);
            double z, s, t;
            const int h = cast(ps, n, d, &z, &s, &t);
            double inten = 0.0;
            if (h >= 0) inten = texture(&ps[h], s, t);
            if (noise) inten += 2.0 * gauss(seed, 100 + frame, idx);
            if (I) I[idx] = to_u8(inten);
            uint16_t zq = (h >= 0) ? to_u16_depth(z) : 0;
            if (hash_unit(seed, 10 + frame, idx) < 0.05) zq = 0; /* 5% holes */
            if (Z) Z[idx] = zq;
            if (hit) hit[idx] = (zq != 0) ? h : -1;
            if (zout) zout[idx] = z;
        }
}

int pdf_synth_pair(uint64_t seed, unsigned rows, unsigned cols, uint8_t *i1, uint16_t *z1,
                   uint8_t *i2, uint16_t *z2, float *gt)
{
    if (rows == 0 || cols == 0) return 1;
    patch_t ps[MAX_PATCHES], ps2[MAX_PATCHES];
    motion_t mot[MAX_PATCHES];
    int n;
    make_scene(seed, rows, cols, ps, mot, &n);
    for (int i = 0; i < n; i++) ps2[i] = moved_patch(&ps[i], &mot[i]);

    const size_t P = (size_t)rows * cols;
    int *hit = NULL;
    double *zz = NULL;
    if (gt) {
        hit = (int *)malloc(P * sizeof(int));
        zz = (double *)malloc(P * sizeof(double));
        if (!hit || !zz) { free(hit); free(zz); return 2; }
    }
    render(seed, 0, ps, n, rows, cols, 0, i1, z1, hit, zz);
    render(seed, 1, ps2, n, rows, cols, 1, i2, z2, NULL, NULL);

    if (gt) {
        const double f = 525.0 * cols / 640.0;
        const double cx = 0.5 * (cols - 1.0), cy = 0.5 * (rows - 1.0);
        for (unsigned u = 0; u < cols; u++)
            for (unsigned v = 0; v < rows; v++) {
                const size_t idx = v + (size_t)u * rows;
                const int h = hit[idx];
                if (h < 0) { gt[idx] = gt[P + idx] = gt[2 * P + idx] = 0.f; continue; }
                vec3 x1 = vscale(v3((u - cx) / f, (v - cy) / f, 1.0), zz[idx]);
                vec3 rel = vsub(x1, ps[h].p0);
                vec3 x2 = vadd(vadd(vrot(rel, mot[h].axis, mot[h].angle), ps[h].p0), mot[h].trans);
                vec3 m = vsub(x2, x1);
                gt[idx] = (float)m.z;
                gt[P + idx] = (float)m.x;
                gt[2 * P + idx] = (float)m.y;
            }
        free(hit);
        free(zz);
    }
    return 0;
}

static float depth_to_float(uint16_t z) { return (float)((double)z / 5000.0); }

int pdf_synth_sequence_float(uint64_t seed, unsigned rows, unsigned cols, unsigned t, float *I,
                             float *Z)
{
    if (rows == 0 || cols == 0) return 1;
    patch_t ps[MAX_PATCHES], pt[MAX_PATCHES];
    motion_t mot[MAX_PATCHES];
    int n;
    make_scene(seed, rows, cols, ps, mot, &n);
    for (int i = 0; i < n; i++) pt[i] = moved_patch_t(&ps[i], &mot[i], (double)t);
    const size_t P = (size_t)rows * cols;
    uint8_t *a = (uint8_t *)malloc(P);
    uint16_t *b = (uint16_t *)malloc(P * sizeof(uint16_t));
    if (!a || !b) { free(a); free(b); return 2; }
    render(seed, t, pt, n, rows, cols, t > 0, a, b, NULL, NULL);
    for (size_t i = 0; i < P; i++) {
        if (I) I[i] = (float)a[i];
        if (Z) Z[i] = depth_to_float(b[i]);
    }
    free(a);
    free(b);
    return 0;
}

int pdf_synth_pair_float(uint64_t seed, unsigned rows, unsigned cols, float *I1, float *Z1,
                         float *I2, float *Z2, float *gt)
{
    const size_t P = (size_t)rows * cols;
    uint8_t *a = (uint8_t *)malloc(2 * P);
    uint16_t *b = (uint16_t *)malloc(2 * P * sizeof(uint16_t));
    if (!a || !b) { free(a); free(b); return 2; }
    int rc = pdf_synth_pair(seed, rows, cols, a, b, a + P, b + P, gt);
    if (rc == 0) {
        for (size_t i = 0; i < P; i++) {
            if (I1) I1[i] = (float)a[i];
            if (Z1) Z1[i] = depth_to_float(b[i]);
            if (I2) I2[i] = (float)a[P + i];
            if (Z2) Z2[i] = depth_to_float(b[P + i]);
        }
    }
    free(a);
    free(b);
    return rc;
}

int pdf_synth_static_float(uint64_t seed, unsigned rows, unsigned cols, float *I, float *Z)
{
    const size_t P = (size_t)rows * cols;
    uint8_t *a = (uint8_t *)malloc(P);
    uint16_t *b = (uint16_t *)malloc(P * sizeof(uint16_t));
    if (!a || !b) { free(a); free(b); return 2; }
    int rc = pdf_synth_pair(seed, rows, cols, a, b, NULL, NULL, NULL);
    if (rc == 0)
        for (size_t i = 0; i < P; i++) {
            if (I) I[i] = (float)a[i];
            if (Z) Z[i] = depth_to_float(b[i]);
        }
    free(a);
    free(b);
    return rc;
}

uint64_t pdf_synth_hash(const void *data, uint64_t nbytes)
{
    const unsigned char *p = (const unsigned char *)data;
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < nbytes; i++) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

int pdf_synth_pairs_float(uint64_t seed0, unsigned n, unsigned rows, unsigned cols, float *I1,
                          float *Z1, float *I2, float *Z2, float *gt, int threads)
{
    const size_t P = (size_t)rows * cols;
    int rc = 0;
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : rc)
    for (long i = 0; i < (long)n; i++)
        rc |= pdf_synth_pair_float(seed0 + (uint64_t)i, rows, cols, I1 ? I1 + i * P : NULL,
                                   Z1 ? Z1 + i * P : NULL, I2 ? I2 + i * P : NULL,
                                   Z2 ? Z2 + i * P : NULL, gt ? gt + 3 * i * P : NULL);
    return rc;
}
