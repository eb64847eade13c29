/*
 * mlp_oracle.c -- FP64 CPU restatement of the learned preconditioner's
 * network and Kaporin loss (see mlp_oracle.h).  TEST INFRASTRUCTURE ONLY.
 */
#include "mlp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kGeluC = 0.7978845608028654;   /* sqrt(2 / pi) */

static double gelu(double z)
{
    return 0.5 * z * (1.0 + tanh(kGeluC * (z + 0.044715 * z * z * z)));
}

static double gelu_grad(double z)
{
    const double t = tanh(kGeluC * (z + 0.044715 * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * kGeluC * (1.0 + 3.0 * 0.044715 * z * z);
}

size_t oracle_mlp_num_params(int F, int H, int nh, int O)
{
    size_t n = (size_t)F * H + H;
    for (int i = 1; i < nh; i++) n += (size_t)H * H + H;
    return n + (size_t)H * O + O;
}

/* layer i: in_i x out_i weights followed by out_i biases */
static void layer_dims(int F, int H, int nh, int O, int i, int *in, int *out)
{
    *in = (i == 0) ? F : H;
    *out = (i == nh) ? O : H;
}

/* pre[i]: B x out_i pre-activations; act[i]: B x out_i activations (act[nh] = y) */
static void forward_all(int F, int H, int nh, int O, int B, const double *params, const double *X, double **pre,
                        double **act)
{
    const double *p = params;
    const double *in_act = X;
    for (int i = 0; i <= nh; i++) {
        int in, out;
        layer_dims(F, H, nh, O, i, &in, &out);
        const double *W = p, *b = p + (size_t)in * out;
        for (int s = 0; s < B; s++)
            for (int j = 0; j < out; j++) {
                double z = b[j];
                for (int k = 0; k < in; k++) z += in_act[(size_t)s * in + k] * W[(size_t)k * out + j];
                pre[i][(size_t)s * out + j] = z;
                act[i][(size_t)s * out + j] = (i == nh) ? z : gelu(z);
            }
        in_act = act[i];
        p += (size_t)in * out + out;
    }
}

void oracle_mlp_forward(int F, int H, int nh, int O, int B, const double *params, const double *X, double *Y)
{
    double **pre = (double **)calloc((size_t)nh + 1, sizeof(double *));
    double **act = (double **)calloc((size_t)nh + 1, sizeof(double *));
    for (int i = 0; i <= nh; i++) {
        int in, out;
        layer_dims(F, H, nh, O, i, &in, &out);
        pre[i] = (double *)malloc(sizeof(double) * (size_t)B * out);
        act[i] = (double *)malloc(sizeof(double) * (size_t)B * out);
    }
    forward_all(F, H, nh, O, B, params, X, pre, act);
    memcpy(Y, act[nh], sizeof(double) * (size_t)B * O);
    for (int i = 0; i <= nh; i++) { free(pre[i]); free(act[i]); }
    free(pre);
    free(act);
}

/* Cholesky of the r x r SPD matrix W in place (lower); returns log det. */
static double chol_logdet(int r, double *W)
{
    double ld = 0.0;
    for (int j = 0; j < r; j++) {
        double d = W[j * r + j];
        for (int k = 0; k < j; k++) d -= W[j * r + k] * W[j * r + k];
        d = sqrt(d);
        W[j * r + j] = d;
        ld += 2.0 * log(d);
        for (int i = j + 1; i < r; i++) {
            double v = W[i * r + j];
            for (int k = 0; k < j; k++) v -= W[i * r + k] * W[j * r + k];
            W[i * r + j] = v / d;
        }
    }
    return ld;
}

double oracle_kaporin(int P, int r, const double *y, const double *As, double *dy)
{
    const int stride = 1 + r;
    const double us = 1.0 / sqrt((double)P);
    double *u = (double *)malloc(sizeof(double) * (size_t)P * (r > 0 ? r : 1));
    double *Au = (double *)calloc((size_t)P * (r > 0 ? r : 1), sizeof(double));
    double *W = (double *)calloc((size_t)(r > 0 ? r * r : 1), sizeof(double));
    double *Winv = (double *)calloc((size_t)(r > 0 ? r * r : 1), sizeof(double));
    double tr = 0.0, ssum = 0.0;
    for (int k = 0; k < P; k++) {
        const double s = y[(size_t)k * stride];
        ssum += s;
        tr += exp(s) * As[(size_t)k * P + k];
        for (int j = 0; j < r; j++) u[(size_t)k * r + j] = y[(size_t)k * stride + 1 + j] * us;
    }
    for (int k = 0; k < P; k++)
        for (int l = 0; l < P; l++) {
            const double a = As[(size_t)k * P + l];
            for (int j = 0; j < r; j++) Au[(size_t)k * r + j] += a * u[(size_t)l * r + j];
        }
    for (int k = 0; k < P; k++)
        for (int j = 0; j < r; j++) tr += u[(size_t)k * r + j] * Au[(size_t)k * r + j];
    for (int a = 0; a < r; a++)
        for (int b = 0; b < r; b++) {
            double v = (a == b) ? 1.0 : 0.0;
            for (int k = 0; k < P; k++) v += exp(-y[(size_t)k * stride]) * u[(size_t)k * r + a] * u[(size_t)k * r + b];
            W[a * r + b] = v;
        }
    const double logdetW = r > 0 ? chol_logdet(r, W) : 0.0;
    const double loss = P * log(tr / P) - ssum - logdetW;
    if (dy) {
        /* W^-1 from the Cholesky factor, column by column */
        for (int c = 0; c < r; c++) {
            double *x = Winv + (size_t)c * r;   /* column c stored as a row (W^-1 is symmetric) */
            for (int i = 0; i < r; i++) {
                double v = (i == c) ? 1.0 : 0.0;
                for (int k = 0; k < i; k++) v -= W[i * r + k] * x[k];
                x[i] = v / W[i * r + i];
            }
            for (int i = r - 1; i >= 0; i--) {
                double v = x[i];
                for (int k = i + 1; k < r; k++) v -= W[k * r + i] * x[k];
                x[i] = v / W[i * r + i];
            }
        }
        const double g = P / tr;
        for (int k = 0; k < P; k++) {
            const double s = y[(size_t)k * stride], es = exp(s), ems = exp(-s);
            double q = 0.0;                       /* u_k W^-1 u_k^T */
            for (int a = 0; a < r; a++)
                for (int b = 0; b < r; b++) q += u[(size_t)k * r + a] * Winv[a * r + b] * u[(size_t)k * r + b];
            dy[(size_t)k * stride] = g * es * As[(size_t)k * P + k] - 1.0 + ems * q;
            for (int j = 0; j < r; j++) {
                double uw = 0.0;
                for (int a = 0; a < r; a++) uw += u[(size_t)k * r + a] * Winv[a * r + j];
                dy[(size_t)k * stride + 1 + j] = (2.0 * g * Au[(size_t)k * r + j] - 2.0 * ems * uw) * us;
            }
        }
    }
    free(u); free(Au); free(W); free(Winv);
    return loss;
}

double oracle_mlp_loss_grad(int F, int H, int nh, int P, int r, int B, const double *params, const double *X,
                            const double *As, double *grad)
{
    const int O = P * (1 + r);
    double **pre = (double **)calloc((size_t)nh + 1, sizeof(double *));
    double **act = (double **)calloc((size_t)nh + 1, sizeof(double *));
    double **dact = (double **)calloc((size_t)nh + 1, sizeof(double *));
    for (int i = 0; i <= nh; i++) {
        int in, out;
        layer_dims(F, H, nh, O, i, &in, &out);
        pre[i] = (double *)malloc(sizeof(double) * (size_t)B * out);
        act[i] = (double *)malloc(sizeof(double) * (size_t)B * out);
        dact[i] = (double *)calloc((size_t)B * out, sizeof(double));
    }
    forward_all(F, H, nh, O, B, params, X, pre, act);
    double loss = 0.0;
    for (int s = 0; s < B; s++)
        loss += oracle_kaporin(P, r, act[nh] + (size_t)s * O, As + (size_t)s * P * P,
                               grad ? dact[nh] + (size_t)s * O : NULL);
    loss /= B;
    if (grad) {
        const size_t np = oracle_mlp_num_params(F, H, nh, O);
        memset(grad, 0, sizeof(double) * np);
        for (size_t e = 0; e < (size_t)B * O; e++) dact[nh][e] /= B;
        /* parameter offsets */
        size_t *off = (size_t *)malloc(sizeof(size_t) * ((size_t)nh + 1));
        size_t o = 0;
        for (int i = 0; i <= nh; i++) {
            int in, out;
            layer_dims(F, H, nh, O, i, &in, &out);
            off[i] = o;
            o += (size_t)in * out + out;
        }
        for (int i = nh; i >= 0; i--) {
            int in, out;
            layer_dims(F, H, nh, O, i, &in, &out);
            /* dz = dact * g'(pre) (output layer is linear) */
            double *dz = (double *)malloc(sizeof(double) * (size_t)B * out);
            for (size_t e = 0; e < (size_t)B * out; e++) dz[e] = (i == nh) ? dact[i][e] : dact[i][e] * gelu_grad(pre[i][e]);
            const double *in_act = (i == 0) ? X : act[i - 1];
            double *gW = grad + off[i], *gb = gW + (size_t)in * out;
            const double *W = params + off[i];
            for (int s = 0; s < B; s++)
                for (int j = 0; j < out; j++) {
                    const double d = dz[(size_t)s * out + j];
                    gb[j] += d;
                    for (int k = 0; k < in; k++) gW[(size_t)k * out + j] += in_act[(size_t)s * in + k] * d;
                }
            if (i > 0)
                for (int s = 0; s < B; s++)
                    for (int k = 0; k < in; k++) {
                        double v = 0.0;
                        for (int j = 0; j < out; j++) v += dz[(size_t)s * out + j] * W[(size_t)k * out + j];
                        dact[i - 1][(size_t)s * in + k] = v;
                    }
            free(dz);
        }
        free(off);
    }
    for (int i = 0; i <= nh; i++) { free(pre[i]); free(act[i]); free(dact[i]); }
    free(pre); free(act); free(dact);
    return loss;
}
