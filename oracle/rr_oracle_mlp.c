/*
 * rr_oracle_mlp.c -- FP64 restatement of the learned guide (TEST INFRASTRUCTURE ONLY).
 *
 * The reference has no learned guide (its guide is the single-lobe grid of
 * GuidingGrid.slang); this network is new.  Definition shared with the kernels:
 *
 *   x (78)  = [ for c in xyz, j in 0..11: sin(2^j pi u_c), cos(2^j pi u_c) ]  (72,
 *             u = (p - bmin) / (bmax - bmin)) ++ n (3) ++ wo (3)
 *   h1 = relu(W0 x + b0)   64        h2 = relu(W1 h1 + b1)   64
 *   h3 = relu(W2 h2 + b2)  64        h4 = relu(W3 h3 + b3)   64
 *   o  = W4 h4 + b4        40 = 8 lobes x (v3, d, e)
 *   mu_k = v_k / sqrt(|v_k|^2 + 1e-12), kappa_k = 0.1 + softplus(d_k) (<= 1000),
 *   pi = softmax(e)
 *   q(w) = sum_k pi_k vMF(w; mu_k, kappa_k),  vMF = k e^{k(mu.w - 1)} / (2 pi (1 - e^{-2k}))
 *   loss = mean_s  wn_s * -log(q_s(w_s) + 1e-20),  wn = w / mean(w)
 *
 * Parameters are one vector: W0 (64x78 row-major, out x in), b0, W1, b1, W2, b2,
 * W3, b3, W4 (40x64), b4 -> 20,136 values.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>

#include "rr_oracle.h"

#define NI 78
#define NH 64
#define NO 40
#define NK 8

static const int kIn[5] = {NI, NH, NH, NH, NH};
static const int kOut[5] = {NH, NH, NH, NH, NO};

static size_t off_w(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += (size_t)kIn[i] * kOut[i] + kOut[i];
    return o;
}
static size_t off_b(int l) { return off_w(l) + (size_t)kIn[l] * kOut[l]; }

uint32_t orc_mlp_param_count(void) { return (uint32_t)off_w(5); }

void orc_mlp_encode(const float pos[3], const float bmin[3], const float bmax[3], const float n[3],
                    const float wo[3], double* x) {
    const double PI = 3.14159265358979323846;
    for (int c = 0; c < 3; ++c) {
        double u = ((double)pos[c] - bmin[c]) / ((double)bmax[c] - bmin[c]);
        for (int j = 0; j < 12; ++j) {
            double a = ldexp(PI * u, j);
            x[c * 24 + j * 2 + 0] = sin(a);
            x[c * 24 + j * 2 + 1] = cos(a);
        }
    }
    for (int k = 0; k < 3; ++k) { x[72 + k] = n[k]; x[75 + k] = wo[k]; }
}

/* acts: 4 x 64 post-ReLU hidden activations; out: 40 raw outputs */
void orc_mlp_forward(const double* P, const double* x, double* acts, double* out) {
    const double* in = x;
    double tmp[NH];
    for (int l = 0; l < 5; ++l) {
        const double* W = P + off_w(l);
        const double* b = P + off_b(l);
        double* dst = (l < 4) ? acts + l * NH : out;
        for (int o = 0; o < kOut[l]; ++o) {
            double s = b[o];
            for (int i = 0; i < kIn[l]; ++i) s += W[(size_t)o * kIn[l] + i] * in[i];
            tmp[o] = s;
        }
        for (int o = 0; o < kOut[l]; ++o) dst[o] = (l < 4) ? (tmp[o] > 0.0 ? tmp[o] : 0.0) : tmp[o];
        in = dst;
    }
}

static double softplus(double d) { return d > 30.0 ? d : log1p(exp(d)); }
static double sigmoid(double d) { return 1.0 / (1.0 + exp(-d)); }

/* lobes: 8 x (mu3, kappa, pi) */
void orc_mlp_head(const double* out, double* lobes) {
    double emax = -INFINITY;
    for (int k = 0; k < NK; ++k) emax = fmax(emax, out[k * 5 + 4]);
    double z = 0.0;
    for (int k = 0; k < NK; ++k) z += exp(out[k * 5 + 4] - emax);
    for (int k = 0; k < NK; ++k) {
        const double* v = out + k * 5;
        double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-12);
        lobes[k * 5 + 0] = v[0] / len;
        lobes[k * 5 + 1] = v[1] / len;
        lobes[k * 5 + 2] = v[2] / len;
        double kap = 0.1 + softplus(v[3]);
        lobes[k * 5 + 3] = kap > 1000.0 ? 1000.0 : kap;
        lobes[k * 5 + 4] = exp(v[4] - emax) / z;
    }
}

static double vmf(double kap, double c) {
    return kap * exp(kap * (c - 1.0)) / (2.0 * 3.14159265358979323846 * (1.0 - exp(-2.0 * kap)));
}

double orc_mlp_pdf(const double* lobes, const float dir[3]) {
    double q = 0.0;
    for (int k = 0; k < NK; ++k) {
        double c = lobes[k * 5] * dir[0] + lobes[k * 5 + 1] * dir[1] + lobes[k * 5 + 2] * dir[2];
        q += lobes[k * 5 + 4] * vmf(lobes[k * 5 + 3], c);
    }
    return q;
}

/* loss and gradient over a batch: x (B x 78), dirs (B x 3), w (B) -> grad (param count) */
double orc_mlp_loss_grad(const double* P, const double* x, const float* dirs, const float* w, uint32_t B,
                         double* grad) {
    uint32_t np = orc_mlp_param_count();
    memset(grad, 0, np * sizeof(double));
    double wmean = 0.0;
    for (uint32_t s = 0; s < B; ++s) wmean += w[s];
    wmean /= (double)B;
    double loss = 0.0;
    double acts[4 * NH], out[NO], lobes[NK * 5], dout[NO], dh[NH], dprev[NH];
    for (uint32_t s = 0; s < B; ++s) {
        const double* xs = x + (size_t)s * NI;
        const float* d = dirs + (size_t)s * 3;
        double ws = w[s] / wmean / (double)B;
        orc_mlp_forward(P, xs, acts, out);
        orc_mlp_head(out, lobes);
        double f[NK], c[NK], q = 0.0;
        for (int k = 0; k < NK; ++k) {
            c[k] = lobes[k * 5] * d[0] + lobes[k * 5 + 1] * d[1] + lobes[k * 5 + 2] * d[2];
            f[k] = vmf(lobes[k * 5 + 3], c[k]);
            q += lobes[k * 5 + 4] * f[k];
        }
        loss += -ws * log(q + 1e-20);
        double dq = -ws / (q + 1e-20);
        for (int k = 0; k < NK; ++k) {
            const double* v = out + k * 5;
            double pik = lobes[k * 5 + 4], kap = lobes[k * 5 + 3];
            /* softmax */
            dout[k * 5 + 4] = dq * pik * (f[k] - q);
            /* kappa = 0.1 + softplus(d), clamped at 1000 (no gradient beyond) */
            double e2 = exp(-2.0 * kap);
            double dlogf = 1.0 / kap - 2.0 * e2 / (1.0 - e2) + (c[k] - 1.0);
            double raw_kap = 0.1 + softplus(v[3]);
            dout[k * 5 + 3] = (raw_kap > 1000.0) ? 0.0 : dq * pik * f[k] * dlogf * sigmoid(v[3]);
            /* mu = v / |v|: d/dv = (I - mu mu^T) g / |v|, g = dq pi f kappa dir */
            double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-12);
            double g[3], mu[3] = {lobes[k * 5], lobes[k * 5 + 1], lobes[k * 5 + 2]};
            for (int j = 0; j < 3; ++j) g[j] = dq * pik * f[k] * kap * d[j];
            double mg = mu[0] * g[0] + mu[1] * g[1] + mu[2] * g[2];
            for (int j = 0; j < 3; ++j) dout[k * 5 + j] = (g[j] - mu[j] * mg) / len;
        }
        /* back through the layers */
        const double* dcur = dout;
        for (int l = 4; l >= 0; --l) {
            const double* W = P + off_w(l);
            const double* in = (l == 0) ? xs : acts + (l - 1) * NH;
            double* gW = grad + off_w(l);
            double* gb = grad + off_b(l);
            for (int o = 0; o < kOut[l]; ++o) {
                gb[o] += dcur[o];
                for (int i = 0; i < kIn[l]; ++i) gW[(size_t)o * kIn[l] + i] += dcur[o] * in[i];
            }
            if (l == 0) break;
            for (int i = 0; i < kIn[l]; ++i) {
                double sacc = 0.0;
                for (int o = 0; o < kOut[l]; ++o) sacc += W[(size_t)o * kIn[l] + i] * dcur[o];
                dh[i] = (in[i] > 0.0) ? sacc : 0.0; /* ReLU' from the post-activation */
            }
            memcpy(dprev, dh, sizeof(dh));
            dcur = dprev;
        }
    }
    return loss;
}

/* Adam (Kingma & Ba 2015) with bias correction; t is the 1-based step */
void orc_adam(double* P, const double* grad, double* m, double* v, uint32_t n, uint32_t t, double lr,
              double b1, double b2, double eps) {
    double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
    for (uint32_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * grad[i];
        v[i] = b2 * v[i] + (1.0 - b2) * grad[i] * grad[i];
        P[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
