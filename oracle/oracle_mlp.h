/*
 * oracle_mlp.h -- FP64 reference of the warm-start MLP (dense layers, tanh
 * on hidden layers, linear output, mean-squared-error loss, Adam).  No
 * reference counterpart exists; the device FP32 kernels are checked
 * against this, and this is checked against finite differences.
 * TEST INFRASTRUCTURE ONLY.
 */
#ifndef SCFT_ORACLE_MLP_H
#define SCFT_ORACLE_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

/* dims[0] = inputs, dims[n_layers] = outputs.  Parameters are packed layer
 * by layer: W_l (dims[l+1] x dims[l], row-major) then b_l (dims[l+1]). */
long orc_mlp_param_count(int n_layers, const int *dims);
void orc_mlp_forward(int n_layers, const int *dims, const double *params,
                     int batch, const double *X, double *Y);
double orc_mlp_loss_grad(int n_layers, const int *dims, const double *params,
                         int batch, const double *X, const double *T,
                         double *grad);
void orc_adam_step(long n, double *params, const double *grad, double *m,
                   double *v, int t, double lr, double b1, double b2,
                   double eps);

#ifdef __cplusplus
}
#endif
#endif
