/*
 * mlp_oracle.c -- CPU oracle of the learned closure network.
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP training and inference
 * kernels; the product never links it).
 *
 * Network (the closure's specification; no reference counterpart):
 *   x in R^7  ->  h1 = relu(W1 x + b1)      W1: 32x7
 *             ->  h2 = relu(W2 h1 + b2)     W2: 32x32
 *             ->  z  = w3 . h2 + b3         w3: 32
 *             ->  y  = softplus(z) = log(1 + e^z)
 * Parameters, flat, in this order: W1 (row-major, [unit][input]), b1, W2
 * ([out][in]), b2, w3, b3 -- MLP_NPARAM = 1345 values.
 * Loss: mean over the batch of (y - t)^2 / 2.
 * Adam: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 *       p -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).
 * Inference as the collide kernel runs it: features and weights rounded
 * to bf16 (round to nearest even), h1 rounded to bf16 before the second
 * layer, biases in float, products accumulated exactly (double here).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define MLP_IN 7
#define MLP_H 32
#define MLP_NPARAM (MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1)

int mlp_nparam(void) { return MLP_NPARAM; }

static double softplus(double z) { return z > 30.0 ? z : log1p(exp(z)); }
static double sigmoid(double z) { return 1.0 / (1.0 + exp(-z)); }

typedef struct {
    const double *W1, *b1, *W2, *b2, *w3, *b3;
} mlp_view;

static mlp_view view(const double *p)
{
    mlp_view v;
    v.W1 = p;
    v.b1 = v.W1 + MLP_H * MLP_IN;
    v.W2 = v.b1 + MLP_H;
    v.b2 = v.W2 + MLP_H * MLP_H;
    v.w3 = v.b2 + MLP_H;
    v.b3 = v.w3 + MLP_H;
    return v;
}

/* one sample: activations out, returns z */
static double forward1(const mlp_view *m, const double *x, double *h1, double *h2)
{
    for (int j = 0; j < MLP_H; ++j) {
        double a = m->b1[j];
        for (int i = 0; i < MLP_IN; ++i) a += m->W1[j * MLP_IN + i] * x[i];
        h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < MLP_H; ++j) {
        double a = m->b2[j];
        for (int i = 0; i < MLP_H; ++i) a += m->W2[j * MLP_H + i] * h1[i];
        h2[j] = a > 0 ? a : 0;
    }
    double z = m->b3[0];
    for (int j = 0; j < MLP_H; ++j) z += m->w3[j] * h2[j];
    return z;
}

void mlp_forward_f64(const double *params, const double *x, int n, double *y)
{
    const mlp_view m = view(params);
    double h1[MLP_H], h2[MLP_H];
    for (int s = 0; s < n; ++s) y[s] = softplus(forward1(&m, x + (long)s * MLP_IN, h1, h2));
}

/* mean loss and its gradient over n samples */
double mlp_loss_grad_f64(const double *params, const double *x, const double *t, int n,
                         double *grad)
{
    const mlp_view m = view(params);
    memset(grad, 0, sizeof(double) * MLP_NPARAM);
    double *gW1 = grad, *gb1 = gW1 + MLP_H * MLP_IN, *gW2 = gb1 + MLP_H,
           *gb2 = gW2 + MLP_H * MLP_H, *gw3 = gb2 + MLP_H, *gb3 = gw3 + MLP_H;
    double loss = 0, h1[MLP_H], h2[MLP_H], d2[MLP_H], d1[MLP_H];
    for (int s = 0; s < n; ++s) {
        const double *xs = x + (long)s * MLP_IN;
        const double z = forward1(&m, xs, h1, h2);
        const double y = softplus(z), r = y - t[s];
        loss += 0.5 * r * r;
        const double dz = r * sigmoid(z) / n;
        gb3[0] += dz;
        for (int j = 0; j < MLP_H; ++j) {
            gw3[j] += dz * h2[j];
            d2[j] = h2[j] > 0 ? dz * m.w3[j] : 0;
        }
        for (int j = 0; j < MLP_H; ++j) {
            gb2[j] += d2[j];
            for (int i = 0; i < MLP_H; ++i) gW2[j * MLP_H + i] += d2[j] * h1[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            double a = 0;
            for (int j = 0; j < MLP_H; ++j) a += m.W2[j * MLP_H + i] * d2[j];
            d1[i] = h1[i] > 0 ? a : 0;
        }
        for (int j = 0; j < MLP_H; ++j) {
            gb1[j] += d1[j];
            for (int i = 0; i < MLP_IN; ++i) gW1[j * MLP_IN + i] += d1[j] * xs[i];
        }
    }
    return loss / n;
}

void mlp_adam_f64(double *params, const double *grad, double *mom, double *vel, int step,
                  double lr, double beta1, double beta2, double eps)
{
    const double c1 = 1.0 - pow(beta1, step), c2 = 1.0 - pow(beta2, step);
    for (int k = 0; k < MLP_NPARAM; ++k) {
        mom[k] = beta1 * mom[k] + (1.0 - beta1) * grad[k];
        vel[k] = beta2 * vel[k] + (1.0 - beta2) * grad[k] * grad[k];
        params[k] -= lr * (mom[k] / c1) / (sqrt(vel[k] / c2) + eps);
    }
}

/* float -> bf16 -> float, round to nearest even (NaN kept quiet) */
float mlp_round_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        if (u & 0x007fffffu) u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    float r;
    memcpy(&r, &u, 4);
    return r;
}

/* z before softplus, as the collide kernel infers it: x, W1, W2, w3
 * rounded to bf16 (from float), h1 rounded to bf16, biases float. */
void mlp_infer_bf16(const float *params, const float *x, int n, double *z_out)
{
    double W1[MLP_H * MLP_IN], W2[MLP_H * MLP_H], w3[MLP_H];
    const float *b1 = params + MLP_H * MLP_IN, *b2 = b1 + MLP_H + MLP_H * MLP_H,
                *b3 = b2 + MLP_H + MLP_H;
    for (int k = 0; k < MLP_H * MLP_IN; ++k) W1[k] = mlp_round_bf16(params[k]);
    for (int k = 0; k < MLP_H * MLP_H; ++k) W2[k] = mlp_round_bf16(b1[MLP_H + k]);
    for (int k = 0; k < MLP_H; ++k) w3[k] = mlp_round_bf16(b2[MLP_H + k]);
    for (int s = 0; s < n; ++s) {
        const float *xs = x + (long)s * MLP_IN;
        double h1[MLP_H], h2[MLP_H];
        for (int j = 0; j < MLP_H; ++j) {
            double a = b1[j];
            for (int i = 0; i < MLP_IN; ++i) a += W1[j * MLP_IN + i] * (double)mlp_round_bf16(xs[i]);
            const float hf = (float)(a > 0 ? a : 0);
            h1[j] = mlp_round_bf16(hf);
        }
        for (int j = 0; j < MLP_H; ++j) {
            double a = b2[j];
            for (int i = 0; i < MLP_H; ++i) a += W2[j * MLP_H + i] * h1[i];
            h2[j] = a > 0 ? a : 0;
        }
        double z = b3[0];
        for (int j = 0; j < MLP_H; ++j) z += w3[j] * h2[j];
        z_out[s] = z;
    }
}
