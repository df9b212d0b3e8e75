/*
 * CPU oracle for the sphere-sweeping stereo per-frame path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm of
 * KAIST-VCLAB/sphere-stereo's per-frame path so that the HIP library can be
 * checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the shipped library never links or calls it.
 *
 * It is compiled twice (see oracle/Makefile):
 *   ORA_F64=0 -> symbols *_f32: fp32 arithmetic in the reference's operation
 *                order (the bit-exact parity target of the HIP kernels);
 *   ORA_F64=1 -> symbols *_f64: the same method evaluated in fp64 (the error
 *                yardstick: results of the fp32 forms are reported against it).
 * Both builds use -ffp-contract=off and no -ffast-math, so every operation is
 * one IEEE-754 round-to-nearest operation in the written order.  The loops
 * over output pixels run on OpenMP threads (every pixel is computed by one
 * thread from inputs no other thread writes, so results do not depend on the
 * thread count); the scatter of ora_reproject_zmin stays serial.
 *
 * Layouts follow the reference: images HWC (BGR channel order, as cv2 reads
 * them), cost volumes planar [N][H][W], guides uint8 [H][W][3].
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   ora_exp          - against libm exp over the whole argument range (<= 2 ulp).
 *   t_unproject/t_project (double-sphere, torch expressions) - golden vectors
 *                      written by importing the reference's utils.py, closed
 *                      forms (pinhole reduction), round trips, homogeneity.
 *   ora_select_camera, ora_sweep - golden vectors from the reference's own
 *                      RGBD_Estimator methods on torch-CPU, plus brute-force and
 *                      invariant checks (reference camera never selected,
 *                      zero cost at the true depth).
 *   ora_isb_down/up/filter - invariants: constant cost preserved, linearity
 *                      and candidate permutation, constant-guide reduction to a
 *                      fixed stencil computed independently in numpy, odd sizes.
 *   ora_argmin_refine - golden vectors from the reference's
 *                      estimate_fisheye_distance, exact-parabola known answers,
 *                      ties, endpoints, constant volumes.
 *   stitcher functions - closed forms and invariants (identity reprojection,
 *                      inpaint of isolated holes, LUT packing round trip, blend
 *                      weights summing to 1, merge of constant maps).
 *   ora_soft_argmin  - the differentiation form's readout (no reference
 *                      counterpart): closed forms (constant cost, a single
 *                      dominant minimum, geometric weights of a linear cost).
 *   ora_resize_linear - PARITY UNPINNED against cv2 (absent here, and its
 *                      float resize varies with the platform's FMA use):
 *                      closed forms (identity size, 2:1 block means, constant
 *                      and ramp images, border clamps) and agreement with
 *                      torch's bilinear interpolation within float tolerance.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORA_F64
#define ORA_F64 0
#endif

#if ORA_F64
typedef double real;
#define SFX(n) n##_f64
#define K(x) (x)
#define RSQRT(x) sqrt(x)
#define RFLOOR(x) floor(x)
#define RFABS(x) fabs(x)
#define RFMIN(a, b) fmin(a, b)
#define RFMAX(a, b) fmax(a, b)
#else
typedef float real;
#define SFX(n) n##_f32
#define K(x) (x##f)
#define RSQRT(x) sqrtf(x)
#define RFLOOR(x) floorf(x)
#define RFABS(x) fabsf(x)
#define RFMIN(a, b) fminf(a, b)
#define RFMAX(a, b) fmaxf(a, b)
#endif

#define PI_REF 3.14159265 /* stitcher.cu:325, a double literal */

/* ------------------------------------------------------------------------ */
/* exp: Cody-Waite reduction by ln2 = C1 + C2 and the degree-6 polynomial of  */
/* the Cephes math library's expf (S. L. Moshier).  The reference calls the   */
/* CUDA intrinsic __expf (isb_filter.cu:87,229,236,241,248) and expf          */
/* (stitcher.cu:389); this build uses one published method everywhere.        */
/* The final scaling by 2^n is one correctly rounded ldexpf, so results in    */
/* the subnormal range keep their value as CUDA's expf keeps it.              */
/* ------------------------------------------------------------------------ */
static real ora_exp(real x)
{
#if ORA_F64
    return exp(x);
#else
    if (!(x >= -104.0f)) {
        if (x != x) return x;
        return 0.0f;
    }
    if (x > 88.7228394f) return INFINITY;
    float n = floorf(x * 1.44269504088896341f + 0.5f);
    float r = x - n * 0.693359375f;
    r = r - n * -2.12194440e-4f;
    float z = r * r;
    float p = 1.9875691500e-4f;
    p = p * r + 1.3981999507e-3f;
    p = p * r + 8.3334519073e-3f;
    p = p * r + 4.1665795894e-2f;
    p = p * r + 1.6666665459e-1f;
    p = p * r + 5.0000001201e-1f;
    p = p * z + r + 1.0f;
    return ldexpf(p, (int)n);
#endif
}

real SFX(ora_expf)(real x) { return ora_exp(x); }

/* ------------------------------------------------------------------------ */
/* Storage rounding of the cost volume (the precision form): vt = 0 keeps    */
/* fp32, vt = 1 rounds to the nearest binary16 value, vt = 2 to the nearest  */
/* bfloat16 value, ties to even, NaN kept.  Applied where a level-0 volume   */
/* value is stored (the sweep output, an unfused last up pass); the coarser  */
/* pyramid levels and all arithmetic stay fp32.  The fp64 build ignores vt:  */
/* it is the unrounded yardstick.                                            */
/* ------------------------------------------------------------------------ */
#if !ORA_F64
static float round_f16(float x)
{
    if (x != x || x == 0.0f || isinf(x)) return x;
    int e;
    frexpf(x, &e); /* |x| = m * 2^e, m in [0.5, 1) */
    int q = e - 11; /* exponent of one unit in the last place of an 11-bit significand */
    if (q < -24) q = -24; /* binary16 subnormals are multiples of 2^-24 */
    float y = ldexpf(rintf(ldexpf(x, -q)), q); /* exact scalings; rintf rounds to nearest even */
    return fabsf(y) > 65504.0f ? copysignf(INFINITY, x) : y;
}

static float round_bf16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        u |= 0x00400000u;
    else
        u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&x, &u, 4);
    return x;
}
#endif

static real vstore(real x, int vt)
{
#if ORA_F64
    (void)vt;
    return x;
#else
    return vt == 1 ? round_f16(x) : (vt == 2 ? round_bf16(x) : x);
#endif
}

void SFX(ora_round_volume)(real *x, long n, int vt)
{
    for (long i = 0; i < n; i++) x[i] = vstore(x[i], vt);
}

/* ------------------------------------------------------------------------ */
/* Double-sphere model as utils.py:57-94 writes it with torch tensors:       */
/* fp32 tensor ops whose Python-double scalar operands (alpha**2, 2*alpha-1, */
/* 1-xi**2, 1-alpha, -w2) are rounded to fp32 once.  Camera record:         */
/* p[0..1] = fl*matching_scale, p[2..3] = principal*matching_scale (the fp32 */
/* products the reference forms), p[4] = xi, p[5] = alpha (Python doubles).  */
/* ------------------------------------------------------------------------ */
typedef struct {
    real fs0, fs1, cs0, cs1;
    real xi, a, a2, am2_1, one_m_xi2, one_m_a, mw2;
} TCam;

static void tcam_init(TCam *c, const double *p)
{
    double xi = p[4], al = p[5];
    c->fs0 = (real)p[0];
    c->fs1 = (real)p[1];
    c->cs0 = (real)p[2];
    c->cs1 = (real)p[3];
    c->xi = (real)xi;
    c->a = (real)al;
    c->a2 = (real)(al * al);
    c->am2_1 = (real)(2 * al - 1);
    c->one_m_xi2 = (real)(1 - xi * xi);
    c->one_m_a = (real)(1 - al);
    /* utils.py:86-90 in Python double */
    double w1 = al > 0.5 ? (1 - al) / al : al / (1 - al);
    double w2 = (w1 + xi) / sqrt(2 * w1 * xi + xi * xi + 1);
    c->mw2 = (real)(-w2);
}

/* utils.py:57-73 */
static void t_unproject(const TCam *c, real u, real v, real out[3], int *valid)
{
    real mx = (u - c->cs0) / c->fs0;
    real my = (v - c->cs1) / c->fs1;
    real r2 = mx * mx + my * my;
    real t = K(1.0) - c->am2_1 * r2;
    real tc = t > 0 ? t : 0; /* torch.clamp(min=0) */
    real mz = (K(1.0) - c->a2 * r2) / ((c->a * RSQRT(tc) + K(1.0)) - c->a);
    real mz2 = mz * mz;
    real k = (mz * c->xi + RSQRT(mz2 + c->one_m_xi2 * r2)) / (mz2 + r2);
    out[0] = k * mx;
    out[1] = k * my;
    out[2] = k * mz - c->xi;
    if (valid) *valid = t >= 0;
}

/* utils.py:75-94.  torch.norm over the last dim is sqrt((x*x + y*y) + z*z). */
static void t_project(const TCam *c, const real p[3], real uv[2], int *valid)
{
    real xy2 = p[0] * p[0] + p[1] * p[1];
    real d1 = RSQRT(xy2 + p[2] * p[2]);
    real cc = c->xi * d1 + p[2];
    real d2 = RSQRT(xy2 + cc * cc);
    real nrm = c->a * d2 + c->one_m_a * cc;
    if (valid) *valid = p[2] > c->mw2 * d1;
    uv[0] = (c->fs0 * p[0]) / nrm + c->cs0;
    uv[1] = (c->fs1 * p[1]) / nrm + c->cs1;
}

void SFX(ora_unproject_torch)(const double *cam, const double *uv, int n, real *out, int32_t *valid)
{
    TCam c;
    tcam_init(&c, cam);
    for (int i = 0; i < n; i++) {
        int v;
        t_unproject(&c, (real)uv[2 * i], (real)uv[2 * i + 1], out + 3 * i, &v);
        if (valid) valid[i] = v;
    }
}

void SFX(ora_project_torch)(const double *cam, const double *pts, int n, real *uv, int32_t *valid)
{
    TCam c;
    tcam_init(&c, cam);
    for (int i = 0; i < n; i++) {
        real p[3] = {(real)pts[3 * i], (real)pts[3 * i + 1], (real)pts[3 * i + 2]};
        int v;
        t_project(&c, p, uv + 2 * i, &v);
        if (valid) valid[i] = v;
    }
}

/* The homogeneous product torch.matmul(cat([p, 1]), rt.T)
 * (depth_estimation.py:103-104, 157-158): row j is
 * fma(p2, R[j][2], fma(p1, R[j][1], p0 * R[j][0])) + R[j][3], the k-ordered
 * multiply-add chain a BLAS micro-kernel evaluates.  rt is 3x4 row-major. */
static void rigid(const real *rt, const real p[3], real q[3])
{
    for (int j = 0; j < 3; j++) {
        const real *r = rt + 4 * j;
#if ORA_F64
        q[j] = fma(p[2], r[2], fma(p[1], r[1], p[0] * r[0])) + r[3];
#else
        q[j] = fmaf(p[2], r[2], fmaf(p[1], r[1], p[0] * r[0])) + r[3];
#endif
    }
}

/* grid_sample's align_corners=False coordinate round trip
 * (depth_estimation.py:116-117, 160-161 then torch's unnormalize). */
static real grid_coord(real u, int size)
{
    real g = ((u + K(0.5)) / (real)size) * K(2.0) - K(1.0);
    return ((g + K(1.0)) * (real)size - K(1.0)) / K(2.0);
}

/* Bilinear tap weights and zero-padding rule of torch's 5-D grid_sample with
 * D=1 and z=0 (depth_estimation.py:166): taps nw, ne, sw, se accumulated in
 * that order starting from 0, out-of-image taps skipped.  img is HWC with
 * nch channels; out receives nch values. */
static void sample_zero(const float *img, int H, int W, int nch, real ix, real iy, real *out)
{
    for (int c = 0; c < nch; c++) out[c] = 0;
    real x0f = RFLOOR(ix), y0f = RFLOOR(iy);
    if (!(x0f >= -1 && x0f <= (real)(W - 1) && y0f >= -1 && y0f <= (real)(H - 1))) return;
    int x0 = (int)x0f, y0 = (int)y0f;
    real x1f = x0f + K(1.0), y1f = y0f + K(1.0);
    real w[4] = {(x1f - ix) * (y1f - iy), (ix - x0f) * (y1f - iy), (x1f - ix) * (iy - y0f),
                 (ix - x0f) * (iy - y0f)};
    int tx[4] = {x0, x0 + 1, x0, x0 + 1}, ty[4] = {y0, y0, y0 + 1, y0 + 1};
    for (int t = 0; t < 4; t++) {
        if (tx[t] < 0 || tx[t] >= W || ty[t] < 0 || ty[t] >= H) continue;
        const float *px = img + ((size_t)ty[t] * W + tx[t]) * nch;
        for (int c = 0; c < nch; c++) out[c] = out[c] + (real)px[c] * w[t];
    }
}

/* RGBD_Estimator.select_camera (depth_estimation.py:83-135) for one reference.
 * cams: C x 6 camera records; rts: C x 12 (cam_k <- reference, 3x4 row-major,
 * the fp32 matrices torch.matmul(torch.inverse(rt_k), rt_ref) gives); masks:
 * C x H x W.  sel receives -1 or the camera index. */
void SFX(ora_select_camera)(int C, int H, int W, int ref, const double *cams, const double *rts,
                            const float *masks, double min_dist, double max_dist, int32_t *sel)
{
    TCam rc, *cc = (TCam *)malloc(sizeof(TCam) * C);
    real *rt = (real *)malloc(sizeof(real) * 12 * C);
    tcam_init(&rc, cams + 6 * ref);
    for (int k = 0; k < C; k++) tcam_init(&cc[k], cams + 6 * k);
    for (int i = 0; i < 12 * C; i++) rt[i] = (real)rts[i];
    real mn = (real)min_dist, mx = (real)max_dist;
    #pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            real ray[3];
            int ref_valid;
            t_unproject(&rc, (real)x, (real)y, ray, &ref_valid);
            int best = -1;
            real maxd = K(1.0);
            real mref = (real)masks[(size_t)ref * H * W + (size_t)y * W + x];
            for (int k = 0; k < C; k++) {
                real pn[3] = {ray[0] * mn, ray[1] * mn, ray[2] * mn};
                real pf[3] = {ray[0] * mx, ray[1] * mx, ray[2] * mx};
                real qn[3], qf[3], uvn[2], uvf[2];
                rigid(rt + 12 * k, pn, qn);
                rigid(rt + 12 * k, pf, qf);
                real nn = RSQRT((qn[0] * qn[0] + qn[1] * qn[1]) + qn[2] * qn[2]);
                real nf = RSQRT((qf[0] * qf[0] + qf[1] * qf[1]) + qf[2] * qf[2]);
                for (int j = 0; j < 3; j++) {
                    qn[j] = qn[j] / nn;
                    qf[j] = qf[j] / nf;
                }
                int vn, vf;
                t_project(&cc[k], qn, uvn, &vn);
                t_project(&cc[k], qf, uvf, &vf);
                real dx = uvn[0] - uvf[0], dy = uvn[1] - uvf[1];
                real disp = RSQRT(dx * dx + dy * dy);
                real mnear, mfar;
                const float *mk = masks + (size_t)k * H * W;
                sample_zero(mk, H, W, 1, grid_coord(uvn[0], W), grid_coord(uvn[1], H), &mnear);
                sample_zero(mk, H, W, 1, grid_coord(uvf[0], W), grid_coord(uvf[1], H), &mfar);
                int ok = disp > maxd && ref_valid && vn && vf && mref >= K(0.9) && mnear >= K(0.9) &&
                         mfar >= K(0.9);
                if (ok) {
                    maxd = disp;
                    best = k;
                }
            }
            sel[(size_t)y * W + x] = best;
        }
    free(cc);
    free(rt);
}

/* Sweep and raw cost (depth_estimation.py:137-175) for one reference.
 * imgs: C x H x W x 3 (float32 BGR, [0,255]); sel: H x W; cands: the N
 * distance candidates 1/linspace(1/min, 1/max, N); cost: N x H x W,
 * cost = min(500, |s0-r0| + |s1-r1| + |s2-r2|), s = 0 where sel < 0. */
void SFX(ora_sweep)(int C, int H, int W, int ref, const float *imgs, const int32_t *sel,
                    const double *cands, int N, const double *cams, const double *rts, real *cost, int vt)
{
    TCam rc, *cc = (TCam *)malloc(sizeof(TCam) * C);
    real *rt = (real *)malloc(sizeof(real) * 12 * C);
    real *cd = (real *)malloc(sizeof(real) * N);
    tcam_init(&rc, cams + 6 * ref);
    for (int k = 0; k < C; k++) tcam_init(&cc[k], cams + 6 * k);
    for (int i = 0; i < 12 * C; i++) rt[i] = (real)rts[i];
    for (int n = 0; n < N; n++) cd[n] = (real)cands[n];
    const size_t HW = (size_t)H * W;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t pix = (size_t)y * W + x;
            real ray[3];
            t_unproject(&rc, (real)x, (real)y, ray, NULL);
            const float *rp = imgs + (size_t)ref * HW * 3 + pix * 3;
            real r0 = rp[0], r1 = rp[1], r2 = rp[2];
            int k = sel[pix];
            for (int n = 0; n < N; n++) {
                real s[3] = {0, 0, 0};
                if (k >= 0) {
                    real p[3] = {cd[n] * ray[0], cd[n] * ray[1], cd[n] * ray[2]};
                    real q[3], uv[2];
                    rigid(rt + 12 * k, p, q);
                    t_project(&cc[k], q, uv, NULL);
                    sample_zero(imgs + (size_t)k * HW * 3, H, W, 3, grid_coord(uv[0], W),
                                grid_coord(uv[1], H), s);
                }
                real c = (RFABS(s[0] - r0) + RFABS(s[1] - r1)) + RFABS(s[2] - r2);
                cost[(size_t)n * HW + pix] = vstore(c > K(500.0) ? K(500.0) : c, vt);
            }
        }
    free(cc);
    free(rt);
    free(cd);
}

/* rgb2yCbCr (utils.py:144-155) and the uint8 cast (depth_estimation.py:227),
 * applied to channel indices 0,1,2 of the BGR data as the reference does. */
void SFX(ora_guide)(const float *img, int HW, uint8_t *guide)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < HW; i++) {
        real c0 = img[3 * i], c1 = img[3 * i + 1], c2 = img[3 * i + 2];
        real y = ((K(16.0) + K(0.1826) * c0) + K(0.6142) * c1) + K(0.062) * c2;
        real cb = ((K(128.0) - K(0.1006) * c0) - K(0.3386) * c1) + K(0.4392) * c2;
        real cr = ((K(128.0) + K(0.4392) * c0) - K(0.3989) * c1) - K(0.0403) * c2;
        y = y < 16 ? 16 : (y > 235 ? 235 : y);
        cb = cb < 16 ? 16 : (cb > 240 ? 240 : cb);
        cr = cr < 16 ? 16 : (cr > 240 ? 240 : cr);
        guide[3 * i] = (uint8_t)y;
        guide[3 * i + 1] = (uint8_t)cb;
        guide[3 * i + 2] = (uint8_t)cr;
    }
}

/* ------------------------------------------------------------------------ */
/* ISB filter (isb_filter.cu, isb_filter.py:72-136), planar volumes.        */
/* ------------------------------------------------------------------------ */
static real gdiff(const uint8_t *a, const uint8_t *b)
{
    return (RFABS((real)a[0] - (real)b[0]) + RFABS((real)a[1] - (real)b[1])) + RFABS((real)a[2] - (real)b[2]);
}

/* guideDownsample2xKernel (isb_filter.cu:44-119) */
void SFX(ora_isb_down)(const uint8_t *gin, const real *cin, int hin, int win, uint8_t *gout, real *cout,
                       int hout, int wout, int N, double var_inv_i, int vt)
{
    static const real base[8] = {K(0.25), K(0.5), K(0.25), K(0.5), K(0.25), K(0.5), K(0.25), K(0.5)};
    const real vi = (real)var_inv_i;
    const size_t pin = (size_t)hin * win, pout = (size_t)hout * wout;
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < hout; y++)
        for (int x = 0; x < wout; x++) {
            int x2 = 2 * x, y2 = 2 * y;
            int xp = x2 + 1 < win ? x2 + 1 : win - 1, yp = y2 + 1 < hin ? y2 + 1 : hin - 1;
            int xm = x2 - 1 >= 0 ? x2 - 1 : 0, ym = y2 - 1 >= 0 ? y2 - 1 : 0;
            /* neighbour order of isb_filter.cu:68-77 */
            size_t nb[8] = {(size_t)ym * win + xm, (size_t)y2 * win + xm, (size_t)yp * win + xm,
                            (size_t)yp * win + x2, (size_t)yp * win + xp, (size_t)y2 * win + xp,
                            (size_t)ym * win + xp, (size_t)ym * win + x2};
            size_t ctr = (size_t)y2 * win + x2;
            const uint8_t *gc = gin + 3 * ctr;
            real w[8], ws = K(1.0);
            real sg[3] = {(real)gc[0], (real)gc[1], (real)gc[2]};
            for (int i = 0; i < 8; i++) {
                const uint8_t *gn = gin + 3 * nb[i];
                real d = gdiff(gc, gn);
                w[i] = base[i] * (ora_exp(-d * d * vi) + K(0.01));
                ws = ws + w[i];
                for (int c = 0; c < 3; c++) sg[c] = sg[c] + (real)gn[c] * w[i];
            }
            real inv = K(1.0) / ws; /* float3 / float: reciprocal-multiply */
            size_t o = (size_t)y * wout + x;
            for (int c = 0; c < 3; c++) gout[3 * o + c] = (uint8_t)(sg[c] * inv);
            for (int z = 0; z < N; z++) {
                const real *cz = cin + (size_t)z * pin;
                real acc = cz[ctr];
                for (int i = 0; i < 8; i++) acc = acc + cz[nb[i]] * w[i];
                cout[(size_t)z * pout + o] = vstore(acc / ws, vt);
            }
        }
}

/* guideUpsample2xKernel (isb_filter.cu:132-330): coarse (gin, cin) merged into
 * fine (gio, cio) in place. */
void SFX(ora_isb_up)(const uint8_t *gin, const real *cin, int hin, int win, uint8_t *gio, real *cio, int hio,
                     int wio, int N, double weight_up, double weight_down, double var_inv_i, int vt)
{
    const real wu = (real)weight_up, wd = (real)weight_down, vi = (real)var_inv_i;
    const size_t pin = (size_t)hin * win, pio = (size_t)hio * wio;
    /* stencil taps as (row, col) indices into the 3x3 coarse neighbourhood */
    static const int t1[9][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 0}, {1, 1}, {1, 2}, {2, 0}, {2, 1}, {2, 2}};
    static const int t2[6][2] = {{1, 1}, {2, 1}, {1, 0}, {2, 0}, {1, 2}, {2, 2}};
    static const int t3[6][2] = {{1, 1}, {1, 2}, {0, 1}, {0, 2}, {2, 1}, {2, 2}};
    static const int t4[4][2] = {{1, 1}, {1, 2}, {2, 1}, {2, 2}};
    static const real b1[9] = {1, 1, 1, 1, 8, 1, 1, 1, 1};
    static const real b23[6] = {8, 8, 1, 1, 1, 1};
    static const real b4[4] = {1, 1, 1, 1};
    const int (*taps[4])[2] = {t1, t2, t3, t4};
    const real *bases[4] = {b1, b23, b23, b4};
    const int ntap[4] = {9, 6, 6, 4};
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < hin; y++)
        for (int x = 0; x < win; x++) {
            int rows[3] = {y - 1 >= 0 ? y - 1 : 0, y, y + 1 < hin ? y + 1 : hin - 1};
            int cols[3] = {x - 1 >= 0 ? x - 1 : 0, x, x + 1 < win ? x + 1 : win - 1};
            int x2 = 2 * x, y2 = 2 * y;
            int xpu = x2 + 1 < wio ? x2 + 1 : wio - 1, ypu = y2 + 1 < hio ? y2 + 1 : hio - 1;
            size_t fine[4] = {(size_t)y2 * wio + x2, (size_t)ypu * wio + x2, (size_t)y2 * wio + xpu,
                              (size_t)ypu * wio + xpu};
            int setY = y2 != ypu, setX = x2 != xpu;
            int doit[4] = {1, setY, setX, setY && setX};
            size_t tapidx[4][9];
            real w[4][9], inv[4];
            uint8_t gnew[4][3];
            for (int s = 0; s < 4; s++) {
                const uint8_t *gf = gio + 3 * fine[s];
                real sg[3] = {(real)gf[0] * wu, (real)gf[1] * wu, (real)gf[2] * wu};
                real ws = wu;
                for (int i = 0; i < ntap[s]; i++) {
                    size_t ci = (size_t)rows[taps[s][i][0]] * win + cols[taps[s][i][1]];
                    tapidx[s][i] = ci;
                    const uint8_t *gn = gin + 3 * ci;
                    real d = gdiff(gn, gf);
                    w[s][i] = bases[s][i] * (wd * (ora_exp(-d * d * vi) + K(0.01)));
                    for (int c = 0; c < 3; c++) sg[c] = sg[c] + (real)gn[c] * w[s][i];
                    ws = ws + w[s][i];
                }
                inv[s] = K(1.0) / ws;
                for (int c = 0; c < 3; c++) gnew[s][c] = (uint8_t)(sg[c] * inv[s]);
            }
            for (int s = 0; s < 4; s++)
                if (doit[s])
                    for (int c = 0; c < 3; c++) gio[3 * fine[s] + c] = gnew[s][c];
            for (int z = 0; z < N; z++) {
                const real *cz = cin + (size_t)z * pin;
                real *fz = cio + (size_t)z * pio;
                for (int s = 0; s < 4; s++) {
                    if (!doit[s]) continue;
                    real acc = wu * fz[fine[s]];
                    for (int i = 0; i < ntap[s]; i++) acc = acc + w[s][i] * cz[tapidx[s][i]];
                    fz[fine[s]] = vstore(acc * inv[s], vt);
                }
            }
        }
}

static int ilog_levels(int W, int H)
{
    /* isb_filter.py:62: int(min(log2(cols), log2(rows)) - 1) */
    double l = log2((double)W) < log2((double)H) ? log2((double)W) : log2((double)H);
    return (int)(l - 1);
}

int SFX(ora_isb_levels)(int W, int H) { return ilog_levels(W, H); }

/* ISB_Filter.apply (isb_filter.py:72-136): guide (H x W x 3) and cost
 * (N x H x W) are filtered in place. Returns the level count, or -1 when the
 * reference itself would fail (fewer than one level).  vt: storage rounding of
 * the level-0 values the last up pass writes (see vstore; the coarser levels
 * stay fp32); round_level0 = 0 leaves them unrounded, as when the last up pass
 * feeds the argmin directly. */
int SFX(ora_isb_filter)(uint8_t *guide, real *cost, int H, int W, int N, double sigma_i, double sigma_s, int vt,
                        int round_level0)
{
    int L = ilog_levels(W, H);
    if (L < 1) return -1;
    uint8_t **g = (uint8_t **)malloc(sizeof(void *) * L);
    real **c = (real **)malloc(sizeof(void *) * L);
    int *hs = (int *)malloc(sizeof(int) * L), *ws = (int *)malloc(sizeof(int) * L);
    g[0] = guide;
    c[0] = cost;
    hs[0] = H;
    ws[0] = W;
    for (int s = 1; s < L; s++) {
        hs[s] = (int)ceil(H / pow(2.0, s));
        ws[s] = (int)ceil(W / pow(2.0, s));
        g[s] = (uint8_t *)calloc((size_t)hs[s] * ws[s] * 3, 1);
        c[s] = (real *)calloc((size_t)hs[s] * ws[s] * N, sizeof(real));
    }
    double var_inv_s = 1 / (2 * sigma_s * sigma_s);
    double var_inv_i = 1 / (2 * sigma_i * sigma_i);
    for (int s = 1; s < L; s++)
        SFX(ora_isb_down)(g[s - 1], c[s - 1], hs[s - 1], ws[s - 1], g[s], c[s], hs[s], ws[s], N, var_inv_i, 0);
    for (int s = L - 2; s >= 0; s--) {
        double dist = pow(2.0, s) - 0.5;
        double wdn = exp(-(dist * dist) * var_inv_s);
        double wup = 1 - wdn;
        SFX(ora_isb_up)(g[s + 1], c[s + 1], hs[s + 1], ws[s + 1], g[s], c[s], hs[s], ws[s], N, wup, wdn, var_inv_i,
                        (s == 0 && round_level0) ? vt : 0);
    }
    for (int s = 1; s < L; s++) {
        free(g[s]);
        free(c[s]);
    }
    free(g);
    free(c);
    free(hs);
    free(ws);
    return L;
}

/* Argmin, parabola refinement and index->distance (depth_estimation.py:180-198).
 * cost: N x HW planar. d0, dN1 = first and last candidates; A = d0/dN1 - 1 as the
 * reference's 0-dim tensor arithmetic forms it. */
void SFX(ora_argmin_refine)(const real *cost, int N, int HW, double d0_, double dN1_, double A_, real *dist,
                            int32_t *idx_out)
{
    const real d0 = (real)d0_, dN1 = (real)dN1_, A = (real)A_;
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < HW; p++) {
        real m = cost[p], M = cost[p];
        int i = 0;
        for (int n = 1; n < N; n++) {
            real v = cost[(size_t)n * HW + p];
            if (v < m) {
                m = v;
                i = n;
            }
            if (v > M) M = v;
        }
        int il = i - 1 < 0 ? 0 : i - 1, ir = i + 1 > N - 1 ? N - 1 : i + 1;
        real L = cost[(size_t)il * HW + p], R = cost[(size_t)ir * HW + p];
        real var = (K(0.5) * (L - R)) / (((L + R) - K(2.0) * m) + K(1e-8));
        var = var < K(-0.5) ? K(-0.5) : (var > K(0.5) ? K(0.5) : var);
        if (i == N - 1 || i == 0) var = 0;
        real idx = (real)i + var;
        if (M == m) idx = (real)(N - 1);
        real d = d0 / (((A * idx) / (real)(N - 1)) + K(1.0));
        if (RFABS(M - m) < K(1e-8)) d = dN1;
        dist[p] = d;
        if (idx_out) idx_out[p] = i;
    }
}

/* Soft-argmin readout of the differentiation form (no counterpart in the
 * reference, whose hard argmin + parabola, depth_estimation.py:180-198, has
 * no useful derivative): iota = sum_n n * softmax_n(-beta * cost_n) with the
 * minimum subtracted first, and the inverse distance inv0 + (inv1 - inv0) *
 * iota / (N - 1), the candidates being uniform in inverse distance
 * (depth_estimation.py:145-146).  cost: N x HW planar. */
void SFX(ora_soft_argmin)(const real *cost, int N, int HW, double beta_, double inv0_, double inv1_, real *invd)
{
    const real beta = (real)beta_, inv0 = (real)inv0_, dinv = (real)(inv1_ - inv0_);
    #pragma omp parallel for schedule(static)
    for (int p = 0; p < HW; p++) {
        real m = cost[p];
        for (int n = 1; n < N; n++) m = RFMIN(m, cost[(size_t)n * HW + p]);
        real z = 0, s = 0;
        for (int n = 0; n < N; n++) {
#if ORA_F64
            real e = exp(-beta * (cost[(size_t)n * HW + p] - m));
#else
            real e = expf(-beta * (cost[(size_t)n * HW + p] - m));
#endif
            z = z + e;
            s = s + (real)n * e;
        }
        invd[p] = inv0 + dinv * ((s / z) / (real)(N - 1));
    }
}

/* ------------------------------------------------------------------------ */
/* Stitcher (stitcher.cu, stitcher.py).  Intrinsics record as               */
/* vectorize_calibration (stitcher.py:35-46): fx, fy, cx, cy, xi, alpha, all */
/* fp32; the CUDA-side double-sphere functions (stitcher.cu:121-168) compute */
/* in fp32 with no clamp in unproject.                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    real fx, fy, cx, cy, xi, a;
} CIntr;

static void cintr_init(CIntr *c, const double *p)
{
    c->fx = (real)p[0];
    c->fy = (real)p[1];
    c->cx = (real)p[2];
    c->cy = (real)p[3];
    c->xi = (real)p[4];
    c->a = (real)p[5];
}

/* stitcher.cu:121-135 */
static void c_unproject(const CIntr *c, real u, real v, real p[3])
{
    real mx = (u - c->cx) / c->fx, my = (v - c->cy) / c->fy;
    real r2 = mx * mx + my * my;
    real mz = (K(1.0) - c->a * c->a * r2) / (c->a * RSQRT(K(1.0) - (K(2.0) * c->a - K(1.0)) * r2) + K(1.0) - c->a);
    real mz2 = mz * mz;
    real k = (mz * c->xi + RSQRT(mz2 + (K(1.0) - c->xi * c->xi) * r2)) / (mz2 + r2);
    p[0] = mx * k;
    p[1] = my * k;
    p[2] = mz * k - c->xi;
}

/* stitcher.cu:140-168; norm3df and hypotf are sqrt of the ordered sum of
 * squares here; float2 / float is a reciprocal multiply (vec_utils.cuh:90-94). */
static void c_project(const CIntr *c, const real p[3], real uv[2], int *valid)
{
    real d1 = RSQRT((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
    real cc = c->xi * d1 + p[2];
    real d2 = RSQRT((p[0] * p[0] + p[1] * p[1]) + cc * cc);
    real nrm = c->a * d2 + (K(1.0) - c->a) * cc;
    if (valid) {
        real w1 = (double)c->a > 0.5 ? (K(1.0) - c->a) / c->a : c->a / (K(1.0) - c->a);
        real w2 = (w1 + c->xi) / RSQRT(K(2.0) * w1 * c->xi + c->xi * c->xi + K(1.0));
        *valid = *valid && (p[2] > -w2 * d1);
    }
    real inv = K(1.0) / nrm;
    uv[0] = (c->fx * p[0]) * inv + c->cx;
    uv[1] = (c->fy * p[1]) * inv + c->cy;
}

void SFX(ora_unproject_cuda)(const double *intr, const double *uv, int n, real *out)
{
    CIntr c;
    cintr_init(&c, intr);
    for (int i = 0; i < n; i++) c_unproject(&c, (real)uv[2 * i], (real)uv[2 * i + 1], out + 3 * i);
}

void SFX(ora_project_cuda)(const double *intr, const double *pts, int n, real *uv, int32_t *valid)
{
    CIntr c;
    cintr_init(&c, intr);
    for (int i = 0; i < n; i++) {
        real p[3] = {(real)pts[3 * i], (real)pts[3 * i + 1], (real)pts[3 * i + 2]};
        int v = 1;
        c_project(&c, p, uv + 2 * i, valid ? &v : NULL);
        if (valid) valid[i] = v;
    }
}

/* __float2int_rn: round half to even, NaN -> 0, saturating to int32 */
static int32_t cvt_rni(real v)
{
    if (v != v) return 0;
#if ORA_F64
    real r = rint(v);
#else
    real r = rintf(v);
#endif
    if (r >= 2147483647.0) return 2147483647;
    if (r <= -2147483648.0) return (int32_t)(-2147483647 - 1);
    return (int32_t)r;
}

/* Bilinear read with floor (__float2int_rd) and no bounds check
 * (stitcher.cu:56-116); callers keep uv inside. */
static real interp_f(const real *m, int cols, real u, real v)
{
    int u1 = (int)RFLOOR(u), v1 = (int)RFLOOR(v), u2 = u1 + 1, v2 = v1 + 1;
    real u1f = (real)u1, u2f = (real)u2, v1f = (real)v1, v2f = (real)v2;
    real w1 = (u2f - u) * (v2f - v), w2 = (u2f - u) * (v - v1f), w3 = (u - u1f) * (v2f - v),
         w4 = (u - u1f) * (v - v1f);
    real p1 = m[(size_t)v1 * cols + u1], p2 = m[(size_t)v2 * cols + u1], p3 = m[(size_t)v1 * cols + u2],
         p4 = m[(size_t)v2 * cols + u2];
    return ((w1 * p1 + w2 * p2) + w3 * p3) + w4 * p4;
}

static void interp_rgb(const uint8_t *img, int cols, real u, real v, real out[3])
{
    int u1 = (int)RFLOOR(u), v1 = (int)RFLOOR(v), u2 = u1 + 1, v2 = v1 + 1;
    real u1f = (real)u1, u2f = (real)u2, v1f = (real)v1, v2f = (real)v2;
    real w1 = (u2f - u) * (v2f - v), w2 = (u2f - u) * (v - v1f), w3 = (u - u1f) * (v2f - v),
         w4 = (u - u1f) * (v - v1f);
    const uint8_t *p1 = img + ((size_t)v1 * cols + u1) * 3, *p2 = img + ((size_t)v2 * cols + u1) * 3,
                  *p3 = img + ((size_t)v1 * cols + u2) * 3, *p4 = img + ((size_t)v2 * cols + u2) * 3;
    for (int c = 0; c < 3; c++)
        out[c] = (((real)p1[c] * w1 + (real)p2[c] * w2) + (real)p3[c] * w3) + (real)p4[c] * w4;
}

/* reprojectDistanceKernel (stitcher.cu:181-204) with an exact z-minimum: the
 * reference's non-atomic read-compare-write, run twice, approximates this.
 * The target index keeps the reference's linear form rint(y)*W + rint(x)
 * (so x just outside a row lands in the neighbouring row, as there); indices
 * outside [0, H*W) are dropped. dout must hold the fill value (1e8) on entry. */
void SFX(ora_reproject_zmin)(const real *din, real *dout, int H, int W, const double *intr, const double *t3)
{
    CIntr c;
    cintr_init(&c, intr);
    real t[3] = {(real)t3[0], (real)t3[1], (real)t3[2]};
    const long long HW = (long long)H * W;
    for (long long i = 0; i < HW; i++) {
        real p[3], uv[2];
        c_unproject(&c, (real)(i % W), (real)(i / W), p);
        real d = din[i];
        real pt[3] = {p[0] * d - t[0], p[1] * d - t[1], p[2] * d - t[2]};
        c_project(&c, pt, uv, NULL);
        real dist = RSQRT((pt[0] * pt[0] + pt[1] * pt[1]) + pt[2] * pt[2]);
        long long o = (long long)cvt_rni(uv[1]) * W + cvt_rni(uv[0]);
        if (o >= 0 && o < HW && dist < dout[o]) dout[o] = dist;
    }
}

/* createInpaintingWeightsKernel (stitcher.cu:212-272).  lut: H x W x 2. */
void SFX(ora_inpaint_lut)(uint8_t *lut, int H, int W, const double *intr, const double *t3, double min_dist,
                          double max_dist)
{
    CIntr c;
    cintr_init(&c, intr);
    real t[3] = {(real)t3[0], (real)t3[1], (real)t3[2]};
    real mn = (real)min_dist, mx = (real)max_dist;
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            real u[3], pc[2], pf[2];
            c_unproject(&c, (real)x, (real)y, u);
            real a[3] = {u[0] * mn - t[0], u[1] * mn - t[1], u[2] * mn - t[2]};
            real b[3] = {u[0] * mx - t[0], u[1] * mx - t[1], u[2] * mx - t[2]};
            c_project(&c, a, pc, NULL);
            c_project(&c, b, pf, NULL);
            real dx = pf[0] - pc[0], dy = pf[1] - pc[1];
            real il = K(1.0) / RSQRT(dx * dx + dy * dy);
            dx = dx * il;
            dy = dy * il;
            real wt[2] = {0, 0};
            int nbx[2] = {0, 0}, nby[2] = {0, 0};
            for (int n = -1; n <= 1; n++)
                for (int m = -1; m <= 1; m++) {
                    if (!(n || m)) continue;
                    real ip = K(1.0) / RSQRT((real)n * (real)n + (real)m * (real)m);
                    real w = ((real)n * ip) * dx + ((real)m * ip) * dy;
                    if (w > wt[1]) {
                        wt[0] = wt[1];
                        nbx[0] = nbx[1];
                        nby[0] = nby[1];
                        wt[1] = w;
                        nbx[1] = n;
                        nby[1] = m;
                    } else if (w > wt[0]) {
                        wt[0] = w;
                        nbx[0] = n;
                        nby[0] = m;
                    }
                }
            size_t o = ((size_t)y * W + x) * 2;
            for (int s = 0; s < 2; s++)
                lut[o + s] = (uint8_t)((((uint8_t)(wt[s] * K(255.0))) & 240) + 4 * (1 + nby[s]) + (1 + nbx[s]));
        }
}

/* One pass of inpaintKernel (stitcher.cu:283-323) in Jacobi form: reads din,
 * writes dout (the reference updates in place within a launch).  A pixel is
 * empty when (double)d >= thr_empty; a neighbour contributes when
 * (double)d <= thr_src and its 4-bit weight is non-zero. */
void SFX(ora_inpaint_step)(const real *din, real *dout, const uint8_t *lut, int H, int W, double thr_empty,
                           double thr_src)
{
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t i = (size_t)y * W + x;
            real d = din[i];
            dout[i] = d;
            if (!((double)d >= thr_empty)) continue;
            if (!(x < W - 1 && y < H - 1 && y >= 1 && x >= 1)) continue;
            real val = 0, wsum = 0;
            for (int s = 0; s < 2; s++) {
                int b = lut[2 * i + s];
                int nx = x + (b & 3) - 1, ny = y + ((b & 12) / 4) - 1;
                real w = (real)(b & 240);
                real sd = din[(size_t)ny * W + nx];
                if ((double)sd <= thr_src && w > 0) {
                    val = val + w * sd;
                    wsum = wsum + w;
                }
            }
            if (wsum > 0) dout[i] = val / wsum;
        }
}

/* Box smoothing of stitcher.py:145-148 and 162-164 (conv2d with a normalised
 * (2r+1)^2 box): evaluated separably, row sums then column sums, then one
 * multiply by the fp32 value 1/(2r+1)^2.  Outside the image the input reads
 * as pad_value (1 for the masks, 0 for the blending weights).  B planes. */
void SFX(ora_box_filter)(const real *in, real *out, int B, int H, int W, int r, double pad_value)
{
    const int Hp = H + 2 * r;
    const real pv = (real)pad_value;
    const real k = K(1.0) / (real)((2 * r + 1) * (2 * r + 1));
    real *rows = (real *)malloc(sizeof(real) * (size_t)Hp * W);
    for (int b = 0; b < B; b++) {
        const real *src = in + (size_t)b * H * W;
        for (int yy = 0; yy < Hp; yy++) {
            int y = yy - r;
            for (int x = 0; x < W; x++) {
                real s = 0;
                for (int j = -r; j <= r; j++) {
                    int xx = x + j;
                    real v = (y < 0 || y >= H || xx < 0 || xx >= W) ? pv : src[(size_t)y * W + xx];
                    s = s + v;
                }
                rows[(size_t)yy * W + x] = s;
            }
        }
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                real s = 0;
                for (int i = 0; i <= 2 * r; i++) s = s + rows[(size_t)(y + i) * W + x];
                out[(size_t)b * H * W + (size_t)y * W + x] = s * k;
            }
    }
    free(rows);
}

/* createBlendingLutKernel (stitcher.cu:338-401).  masks: R x H x W (already
 * box-smoothed); sampling: R x PH x PW x 2; weights: R x PH x PW (before the
 * box smoothing of stitcher.py:162-164).  rot: R x 9, the fp32 inverse
 * rotations; trans: R x 3.  cosf/sinf are the fp32 roundings of the fp64
 * cos/sin. */
void SFX(ora_blend_lut)(real *sampling, real *weights, const real *masks, int R, int H, int W, int PH, int PW,
                        const double *intr, const double *rot, const double *trans, double min_dist,
                        double max_dist)
{
    if (R > 16) return; /* the per-pixel weights live on the stack */
    CIntr *c = (CIntr *)malloc(sizeof(CIntr) * R);
    for (int r = 0; r < R; r++) cintr_init(&c[r], intr + 6 * r);
    real mn = (real)min_dist, mx = (real)max_dist;
    const size_t P = (size_t)PH * PW;
    #pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < PH; y++)
        for (int x = 0; x < PW; x++) {
            real bw[16];
            size_t idx = (size_t)y * PW + x;
            real phi = (real)(((double)((float)y + 0.5f) * PI_REF) / PH - PI_REF / 2.0);
            real theta = (real)(((double)(((float)x + 0.5f) * 2.0f) * PI_REF) / PW + PI_REF);
            real cp = (real)cos((double)phi), sp = (real)sin((double)phi);
            real ct = (real)cos((double)theta), st = (real)sin((double)theta);
            real unit[3] = {cp * st, sp, cp * ct};
            real bsum = 0;
            for (int r = 0; r < R; r++) {
                real rm[9];
                for (int i = 0; i < 9; i++) rm[i] = (real)rot[9 * r + i];
                real uf[3];
                for (int j = 0; j < 3; j++) uf[j] = (rm[3 * j] * unit[0] + rm[3 * j + 1] * unit[1]) + rm[3 * j + 2] * unit[2];
                int valid = 1;
                real uv[2];
                c_project(&c[r], uf, uv, &valid);
                uv[0] = RFMIN(RFMAX(uv[0], K(0.1)), (real)W - K(1.1));
                uv[1] = RFMIN(RFMAX(uv[1], K(0.1)), (real)H - K(1.1));
                sampling[((size_t)r * P + idx) * 2] = uv[0];
                sampling[((size_t)r * P + idx) * 2 + 1] = uv[1];
                bw[r] = K(1e-8);
                real t[3] = {(real)trans[3 * r], (real)trans[3 * r + 1], (real)trans[3 * r + 2]};
                real a[3] = {uf[0] * mn - t[0], uf[1] * mn - t[1], uf[2] * mn - t[2]};
                real b[3] = {uf[0] * mx - t[0], uf[1] * mx - t[1], uf[2] * mx - t[2]};
                real pn[2], pf[2];
                c_project(&c[r], a, pn, &valid);
                pn[0] = RFMIN(RFMAX(pn[0], K(0.1)), (real)W - K(1.1));
                pn[1] = (real)(r * H) + RFMIN(RFMAX(pn[1], K(0.1)), (real)H - K(1.1));
                c_project(&c[r], b, pf, &valid);
                pf[0] = RFMIN(RFMAX(pf[0], K(0.1)), (real)W - K(1.1));
                pf[1] = (real)(r * H) + RFMIN(RFMAX(pf[1], K(0.1)), (real)H - K(1.1));
                if (valid && (double)interp_f(masks, W, pn[0], pn[1]) > 0.99 &&
                    (double)interp_f(masks, W, pf[0], pf[1]) > 0.99) {
                    real dx = pf[0] - pn[0], dy = pf[1] - pn[1];
                    real ds = RSQRT(dx * dx + dy * dy);
                    bw[r] = ora_exp((real)((double)(-ds * ds) / (1e-4 * H * W)));
                }
                bsum = bsum + bw[r];
            }
            for (int r = 0; r < R; r++) weights[(size_t)r * P + idx] = bw[r] / bsum;
        }
    free(c);
}

/* blending_weights /= sum over references (stitcher.py:164) */
void SFX(ora_normalize_weights)(real *w, int R, int P)
{
    for (int p = 0; p < P; p++) {
        real s = 0;
        for (int r = 0; r < R; r++) s = s + w[(size_t)r * P + p];
        for (int r = 0; r < R; r++) w[(size_t)r * P + p] = w[(size_t)r * P + p] / s;
    }
}

/* mergeRGBDPanoramaKernel (stitcher.cu:423-474).  reproj, dists: R x H x W;
 * imgs: R x Hs x Ws x 3 uint8. */
void SFX(ora_merge)(const real *sampling, const real *weights, const real *reproj, const real *dists,
                    const uint8_t *imgs, int Hs, int Ws, const double *trans, const double *intr, int R, int H, int W,
                    int PH, int PW, real *dist_out, uint8_t *rgb_out)
{
    CIntr *c = (CIntr *)malloc(sizeof(CIntr) * R);
    for (int r = 0; r < R; r++) cintr_init(&c[r], intr + 6 * r);
    const size_t P = (size_t)PH * PW;
    const real rx = (real)Ws / (real)W, ry = (real)Hs / (real)H;
    #pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < P; idx++) {
        real inv_sum = 0, rgb[3] = {0, 0, 0};
        for (int r = 0; r < R; r++) {
            real bw = weights[(size_t)r * P + idx];
            real u = sampling[((size_t)r * P + idx) * 2], v = sampling[((size_t)r * P + idx) * 2 + 1];
            real vo = v + (real)(r * H);
            real rd = interp_f(reproj, W, u, vo);
            inv_sum = inv_sum + (bw * K(1.0)) / rd;
            real dc = interp_f(dists, W, u, vo);
            dc = RFMIN(rd, dc);
            real p[3], uv[2];
            c_unproject(&c[r], u, v, p);
            real pt[3] = {p[0] * dc + (real)trans[3 * r], p[1] * dc + (real)trans[3 * r + 1],
                          p[2] * dc + (real)trans[3 * r + 2]};
            c_project(&c[r], pt, uv, NULL);
            uv[0] = uv[0] * rx;
            uv[1] = uv[1] * ry;
            uv[0] = RFMIN(RFMAX(uv[0], K(0.1)), (real)Ws - K(1.1));
            uv[1] = RFMIN(RFMAX(uv[1], K(0.1)), (real)Hs - K(1.1));
            uv[1] = uv[1] + (real)(r * Hs);
            real col[3];
            interp_rgb(imgs, Ws, uv[0], uv[1], col);
            for (int k = 0; k < 3; k++) rgb[k] = rgb[k] + col[k] * bw;
        }
        dist_out[idx] = K(1.0) / inv_sum;
        for (int k = 0; k < 3; k++) rgb_out[3 * idx + k] = (uint8_t)rgb[k];
    }
    free(c);
}

/* ------------------------------------------------------------------------ */
/* Input stage (utils.py:157-204): the image as float32 in [0, 255], then     */
/* cv2.resize(image, (Wd, Hd), cv2.INTER_AREA), whose flag lands in the dst   */
/* slot, i.e. INTER_LINEAR: for each destination coordinate d, f = (float)    */
/* ((d + 0.5) * (1 / (D / S)) - 0.5) in double, s = floor(f), f -= s, clamped */
/* to the first / last sample (f = 0 there), weights 1 - f and f; then a      */
/* horizontal pass and a vertical pass, each a * w0 + b * w1.                 */
/* src: Hs x Ws x C float32 (already converted), dst: Hd x Wd x C.            */
/* ------------------------------------------------------------------------ */
static void resize_axis(int d, int S, int D, int *s0, int *s1, real *w0, real *w1)
{
    double scale = 1.0 / ((double)D / (double)S);
    real f = (real)((d + 0.5) * scale - 0.5);
    int s = (int)RFLOOR(f);
    f = f - (real)s;
    if (s < 0) {
        f = 0;
        s = 0;
    }
    if (s >= S - 1) {
        f = 0;
        s = S - 1;
    }
    *s0 = s;
    *s1 = s + 1 < S ? s + 1 : S - 1;
    *w0 = K(1.0) - f;
    *w1 = f;
}

void SFX(ora_resize_linear)(const float *src, int Hs, int Ws, int C, real *dst, int Hd, int Wd)
{
    for (int dy = 0; dy < Hd; dy++) {
        int y0, y1;
        real b0, b1;
        resize_axis(dy, Hs, Hd, &y0, &y1, &b0, &b1);
        for (int dx = 0; dx < Wd; dx++) {
            int x0, x1;
            real a0, a1;
            resize_axis(dx, Ws, Wd, &x0, &x1, &a0, &a1);
            for (int c = 0; c < C; c++) {
                real h0 = (real)src[((size_t)y0 * Ws + x0) * C + c] * a0 + (real)src[((size_t)y0 * Ws + x1) * C + c] * a1;
                real h1 = (real)src[((size_t)y1 * Ws + x0) * C + c] * a0 + (real)src[((size_t)y1 * Ws + x1) * C + c] * a1;
                dst[((size_t)dy * Wd + dx) * C + c] = h0 * b0 + h1 * b1;
            }
        }
    }
}
