/*
 * CPU oracle -- test infrastructure only: the path-guiding network.
 *
 * bajo has no learned guiding (its renderer samples the BSDF only,
 * rt/gpu/path_shading.mojo:463-535, rt/shading.mojo:92-118), so this file
 * restates the library's own definition from its documented formulas:
 *  - encoding: 9 inputs x 12 frequencies, (sin, cos)(pi 2^f x): frequencies
 *    0, 4, 8 as sincos_turns(frac(x 2^(f-1))) in FP32 (the renderer's sin/cos
 *    series), the others by FP32 double-angle steps;
 *  - network: 216 -> 64 -> 64 -> 64 -> 40, ReLU, evaluated in FP64; with
 *    `emulate` the operands are rounded to BF16 (round to nearest even)
 *    where the device rounds them (features, weights, hidden activations,
 *    output and hidden deltas), so device and oracle differ only by the
 *    device's FP32 accumulation order;
 *  - output: 8 von Mises-Fisher lobes (mean u/|u|, kappa = e^clamp(s,-4,7),
 *    softmax weights); loss -w ln q(dir) normalized by the batch's sum of w;
 *    analytic gradient by back-propagation in FP64;
 *  - mixture decode / density / sampling restated operation for operation in
 *    FP32 (the device's exp / log series), for bit comparisons.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc_core.h"

void orc_sincos_turns(float u, float *s, float *c);

#define G_IN 9
#define G_FREQ 12
#define G_FEAT 216
#define G_H 64
#define G_OUT 40
#define G_LOBES 8
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + G_H * G_FEAT)
#define OFF_W2 (OFF_B1 + G_H)
#define OFF_B2 (OFF_W2 + G_H * G_H)
#define OFF_W3 (OFF_B2 + G_H)
#define OFF_B3 (OFF_W3 + G_H * G_H)
#define OFF_W4 (OFF_B3 + G_H)
#define OFF_B4 (OFF_W4 + G_OUT * G_H)
#define G_PARAMS (OFF_B4 + G_OUT)
#define KAPPA_LO (-4.0f)
#define KAPPA_HI 7.0f

/* FP32 -> BF16 -> FP32, round to nearest even (finite inputs). */
float orc_bf16_round(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static double rnd(double x, int emulate) { return emulate ? (double)orc_bf16_round((float)x) : x; }

/* Features of one sample as the device computes them: frequencies 0, 4 and 8
 * from the FP32 series of sin/cos(2 pi frac(x 2^(f-1))), the three after each
 * by the double-angle formulas in FP32. */
void orc_guide_encode(const float *x, float *feat) {
    for (int d = 0; d < G_IN; ++d)
        for (int b = 0; b < 3; ++b) {
            float s[4], c[4];
            float y = x[d] * ldexpf(1.0f, 4 * b - 1);
            float u = y - floorf(y);
            orc_sincos_turns(u, &s[0], &c[0]);
            for (int t = 1; t < 4; ++t) {
                s[t] = 2.0f * s[t - 1] * c[t - 1];
                c[t] = (c[t - 1] - s[t - 1]) * (c[t - 1] + s[t - 1]);
            }
            for (int t = 0; t < 4; ++t) {
                feat[24 * d + 2 * (4 * b + t)] = s[t];
                feat[24 * d + 2 * (4 * b + t) + 1] = c[t];
            }
        }
}

typedef struct {
    double feat[G_FEAT], h1[G_H], h2[G_H], h3[G_H], o[G_OUT];
} orc_acts;

static void forward_one(const double *p, const float *x, int emulate, orc_acts *a) {
    float f32[G_FEAT];
    orc_guide_encode(x, f32);
    for (int k = 0; k < G_FEAT; ++k) a->feat[k] = rnd(f32[k], emulate);
    const double *in = a->feat;
    int fan = G_FEAT;
    double *outs[3] = {a->h1, a->h2, a->h3};
    const int woff[3] = {OFF_W1, OFF_W2, OFF_W3}, boff[3] = {OFF_B1, OFF_B2, OFF_B3};
    for (int l = 0; l < 3; ++l) {
        for (int j = 0; j < G_H; ++j) {
            double z = 0.0;
            for (int k = 0; k < fan; ++k) z += rnd(p[woff[l] + j * fan + k], emulate) * in[k];
            z += emulate ? (double)(float)p[boff[l] + j] : p[boff[l] + j];
            outs[l][j] = rnd(z > 0.0 ? z : 0.0, emulate);
        }
        in = outs[l];
        fan = G_H;
    }
    for (int j = 0; j < G_OUT; ++j) {
        double z = 0.0;
        for (int k = 0; k < G_H; ++k) z += rnd(p[OFF_W4 + j * G_H + k], emulate) * a->h3[k];
        a->o[j] = z + (emulate ? (double)(float)p[OFF_B4 + j] : p[OFF_B4 + j]);
    }
}

void orc_guide_forward(const double *params, const float *inputs, uint64_t n, int emulate, double *out) {
    orc_acts a;
    for (uint64_t r = 0; r < n; ++r) {
        forward_one(params, inputs + G_IN * r, emulate, &a);
        memcpy(out + G_OUT * r, a.o, sizeof(a.o));
    }
}

/* -w ln q(dir) and its gradient with respect to the 40 outputs (FP64, log
 * space).  Returns the weighted loss. */
static double output_gradient(const double *o, const double *dir, double w, double *d) {
    double amax = o[4];
    for (int j = 1; j < G_LOBES; ++j) amax = fmax(amax, o[5 * j + 4]);
    double se = 0.0;
    for (int j = 0; j < G_LOBES; ++j) se += exp(o[5 * j + 4] - amax);
    const double lse = amax + log(se);
    double lc[G_LOBES], lp[G_LOBES], kap[G_LOBES], mdw[G_LOBES], len[G_LOBES], e2[G_LOBES], mu[G_LOBES][3];
    int inside[G_LOBES];
    double cmax = -INFINITY;
    for (int j = 0; j < G_LOBES; ++j) {
        const double *oj = o + 5 * j;
        len[j] = sqrt(oj[0] * oj[0] + oj[1] * oj[1] + oj[2] * oj[2]);
        for (int a = 0; a < 3; ++a) mu[j][a] = len[j] > 1e-12 ? oj[a] / len[j] : (a == 2 ? 1.0 : 0.0);
        inside[j] = oj[3] > KAPPA_LO && oj[3] < KAPPA_HI;
        const double s = fmin(fmax(oj[3], KAPPA_LO), KAPPA_HI);
        kap[j] = exp(s);
        e2[j] = exp(-2.0 * kap[j]);
        mdw[j] = mu[j][0] * dir[0] + mu[j][1] * dir[1] + mu[j][2] * dir[2];
        lp[j] = oj[4] - lse;
        lc[j] = lp[j] + (s - log(2.0 * M_PI) - log1p(-e2[j])) + kap[j] * (mdw[j] - 1.0);
        cmax = fmax(cmax, lc[j]);
    }
    double sc = 0.0;
    for (int j = 0; j < G_LOBES; ++j) sc += exp(lc[j] - cmax);
    const double lq = cmax + log(sc);
    for (int j = 0; j < G_LOBES; ++j) {
        const double r = exp(lc[j] - lq), pi = exp(lp[j]);
        d[5 * j + 4] = w * (pi - r);
        const double dlk = 1.0 / kap[j] - 2.0 * e2[j] / (1.0 - e2[j]) + (mdw[j] - 1.0);
        d[5 * j + 3] = inside[j] ? -w * r * dlk * kap[j] : 0.0;
        const double k = len[j] > 1e-12 ? -w * r * kap[j] / len[j] : 0.0;
        for (int a = 0; a < 3; ++a) d[5 * j + a] = k * (dir[a] - mdw[j] * mu[j][a]);
    }
    return -w * lq;
}

/* Normalized loss sum_i -w_i ln q_i / sum_i w_i of samples f32[13 n]
 * (inputs 9, direction 3, weight 1) and, when `grad` is non-NULL, its
 * gradient (FP64 back-propagation; `emulate` rounds as the device does). */
double orc_guide_loss_grad(const double *p, const float *samples, uint64_t n, int emulate, double *grad) {
    orc_acts a;
    double wsum = 0.0, lsum = 0.0;
    if (grad) memset(grad, 0, sizeof(double) * G_PARAMS);
    for (uint64_t r = 0; r < n; ++r) {
        const float *s = samples + 13 * r;
        const double w = s[12];
        if (!(w > 0.0)) continue;
        forward_one(p, s, emulate, &a);
        const double dir[3] = {s[9], s[10], s[11]};
        double d_o[G_OUT];
        lsum += output_gradient(a.o, dir, w, d_o);
        wsum += w;
        if (!grad) continue;
        for (int j = 0; j < G_OUT; ++j) d_o[j] = rnd(d_o[j], emulate);
        /* layer 4 */
        double dh3[G_H] = {0}, dh2[G_H] = {0}, dh1[G_H] = {0};
        for (int j = 0; j < G_OUT; ++j) {
            for (int k = 0; k < G_H; ++k) {
                grad[OFF_W4 + j * G_H + k] += d_o[j] * a.h3[k];
                dh3[k] += d_o[j] * rnd(p[OFF_W4 + j * G_H + k], emulate);
            }
            grad[OFF_B4 + j] += d_o[j];
        }
        for (int k = 0; k < G_H; ++k) dh3[k] = a.h3[k] != 0.0 ? rnd(dh3[k], emulate) : 0.0;
        for (int j = 0; j < G_H; ++j) {
            for (int k = 0; k < G_H; ++k) {
                grad[OFF_W3 + j * G_H + k] += dh3[j] * a.h2[k];
                dh2[k] += dh3[j] * rnd(p[OFF_W3 + j * G_H + k], emulate);
            }
            grad[OFF_B3 + j] += dh3[j];
        }
        for (int k = 0; k < G_H; ++k) dh2[k] = a.h2[k] != 0.0 ? rnd(dh2[k], emulate) : 0.0;
        for (int j = 0; j < G_H; ++j) {
            for (int k = 0; k < G_H; ++k) {
                grad[OFF_W2 + j * G_H + k] += dh2[j] * a.h1[k];
                dh1[k] += dh2[j] * rnd(p[OFF_W2 + j * G_H + k], emulate);
            }
            grad[OFF_B2 + j] += dh2[j];
        }
        for (int k = 0; k < G_H; ++k) dh1[k] = a.h1[k] != 0.0 ? rnd(dh1[k], emulate) : 0.0;
        for (int j = 0; j < G_H; ++j) {
            for (int k = 0; k < G_FEAT; ++k) grad[OFF_W1 + j * G_FEAT + k] += dh1[j] * a.feat[k];
            grad[OFF_B1 + j] += dh1[j];
        }
    }
    if (grad && wsum > 0.0)
        for (int k = 0; k < G_PARAMS; ++k) grad[k] /= wsum;
    return wsum > 0.0 ? lsum / wsum : 0.0;
}

/* One Adam step (Kingma & Ba 2015) in FP64: t is the step number (1-based). */
void orc_guide_adam(double *p, double *m, double *v, const double *g, int t, double lr, double b1, double b2,
                    double eps) {
    const double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
    for (int k = 0; k < G_PARAMS; ++k) {
        m[k] = b1 * m[k] + (1.0 - b1) * g[k];
        v[k] = b2 * v[k] + (1.0 - b2) * g[k] * g[k];
        p[k] -= lr * (m[k] / c1) / (sqrt(v[k] / c2) + eps);
    }
}

/* ---- mixture in FP32, operation for operation as the device -------------------- */
float orc_log2_pos(float x);

static float exp_f(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 87.0f) x = 87.0f;
    const float t = x * 1.44269504088896340736f;
    const float n = rintf(t);
    const float y = (t - n) * 0.693147180559945309417f;
    float p = fmaf(y, 2.48015873015873015873e-5f, 1.98412698412698412698e-4f);
    p = fmaf(y, p, 1.38888888888888888889e-3f);
    p = fmaf(y, p, 8.33333333333333333333e-3f);
    p = fmaf(y, p, 4.16666666666666666667e-2f);
    p = fmaf(y, p, 1.66666666666666666667e-1f);
    p = fmaf(y, p, 0.5f);
    p = fmaf(y, p, 1.0f);
    p = fmaf(y, p, 1.0f);
    uint32_t b = (uint32_t)((int)n + 127) << 23;
    float s;
    memcpy(&s, &b, 4);
    return p * s;
}
static float ln_f(float x) { return orc_log2_pos(x) * 0.693147180559945309417f; }

typedef struct { orc_v3 mu; float kappa, log_norm; } g_lobe;
typedef struct { g_lobe lobe[G_LOBES]; float weight[G_LOBES]; } g_mix;

static g_lobe decode_lobe(const float *o) {
    g_lobe l;
    const float len2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    if (len2 > 1.0e-24f) {
        const float inv = 1.0f / sqrtf(len2);
        l.mu = orc_v(o[0] * inv, o[1] * inv, o[2] * inv);
    } else {
        l.mu = orc_v(0.0f, 0.0f, 1.0f);
    }
    const float s = orc_fmin(orc_fmax(o[3], KAPPA_LO), KAPPA_HI);
    l.kappa = exp_f(s);
    const float e2 = exp_f(-2.0f * l.kappa);
    l.log_norm = s - 1.83787706640934548356f - ln_f(1.0f - e2);
    return l;
}

static void decode(const float *o, g_mix *m) {
    float amax = o[4];
    for (int j = 1; j < G_LOBES; ++j) amax = orc_fmax(amax, o[5 * j + 4]);
    float sum = 0.0f;
    for (int j = 0; j < G_LOBES; ++j) {
        m->lobe[j] = decode_lobe(o + 5 * j);
        m->weight[j] = exp_f(o[5 * j + 4] - amax);
        sum += m->weight[j];
    }
    const float inv = 1.0f / sum;
    for (int j = 0; j < G_LOBES; ++j) m->weight[j] *= inv;
}

static float vmf_pdf(const g_lobe *l, orc_v3 w) { return exp_f(fmaf(l->kappa, orc_dot(l->mu, w) - 1.0f, l->log_norm)); }

static float mix_pdf(const g_mix *m, orc_v3 w) {
    float p = 0.0f;
    for (int j = 0; j < G_LOBES; ++j) p = fmaf(m->weight[j], vmf_pdf(&m->lobe[j], w), p);
    return p;
}

static orc_v3 vmf_sample(const g_lobe *l, float u1, float u2) {
    const float e2 = exp_f(-2.0f * l->kappa);
    float w = 1.0f + ln_f(fmaf(1.0f - u1, e2, u1)) / l->kappa;
    w = orc_fmin(orc_fmax(w, -1.0f), 1.0f);
    const float r = sqrtf(orc_fmax(1.0f - w * w, 0.0f));
    float sp, cp;
    orc_sincos_turns(u2, &sp, &cp);
    /* frame of Duff et al. 2017 around mu */
    const orc_v3 n = l->mu;
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    const orc_v3 t = orc_v(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    const orc_v3 bt = orc_v(b, sign + n.y * n.y * a, -n.y);
    const orc_v3 d = orc_add(orc_add(orc_scale(t, r * cp), orc_scale(bt, r * sp)), orc_scale(n, w));
    return orc_normalize(d);
}

/* Draws a direction of the mixture with outputs o[40] from (u_lobe, u1, u2)
 * and returns its density; `pdf_at` (may be NULL) evaluates the density at
 * a given direction instead of drawing. */
float orc_guide_mixture(const float *o, const float *u, float *dir, const float *pdf_at) {
    g_mix m;
    decode(o, &m);
    if (pdf_at) return mix_pdf(&m, orc_v(pdf_at[0], pdf_at[1], pdf_at[2]));
    int sel = G_LOBES - 1;
    float c = 0.0f;
    for (int j = 0; j < G_LOBES - 1; ++j) {
        c += m.weight[j];
        if (u[0] < c) { sel = j; break; }
    }
    const orc_v3 d = vmf_sample(&m.lobe[sel], u[1], u[2]);
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
    return mix_pdf(&m, d);
}

/* One vMF lobe in FP64: density at cos(angle to the mean) = t, and the
 * cos of a sample from u (Jakob 2012's inversion). */
double orc_vmf_pdf64(double kappa, double t) {
    return kappa / (2.0 * M_PI * (-expm1(-2.0 * kappa))) * exp(kappa * (t - 1.0));
}
double orc_vmf_sample_cos64(double kappa, double u) { return 1.0 + log(u + (1.0 - u) * exp(-2.0 * kappa)) / kappa; }
