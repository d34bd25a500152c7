/* fbse_mlp.h -- learned warm start for the filter (gfx950, BF16 MFMA).
 *
 * No counterpart in the reference: FilterBSE starts every filter from ran2
 * noise (init.c:1207-1305) or from states read back from disk (inputPsiFilt,
 * init.c:1168-1205).  This module adds a small multilayer perceptron that maps
 * per-grid-point features of a dot (distances to the nearest atoms of each
 * species, the local potential, the radial coordinate) to the log-density of
 * its band-edge states; fbse_warm_start_dev turns a predicted density into
 * start states in the reference psi_rank layout (an envelope sqrt(rho) on the
 * ran2 start states), so the plain random-start filter stays the fallback and
 * the measure of the result.
 *
 * Network: n_in -> width -> width -> width -> n_out, SiLU between layers, a
 * linear output; BF16 weights and activations, FP32 accumulation (MFMA
 * v_mfma_f32_16x16x16_bf16), FP32 master weights with Adam.  Training and
 * inference kernels are the module's own (no library GEMMs).
 *
 * Memory layout: feature matrix x is point-major, n x k_in BF16 (k_in = n_in
 * rounded up to a multiple of 16, padding columns zero); outputs y are n x
 * n_out FP32.  Parameters (fbse_mlp_get_params / set_params, FP32) are packed
 * W1 [n_in x width] (row-major: W1[i*width + j] multiplies x_i into unit j),
 * b1 [width], W2 [width x width], b2, W3 [width x width], b3, W4
 * [width x n_out], b4.
 */
#ifndef FBSE_MLP_H
#define FBSE_MLP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fbse_mlp fbse_mlp;
struct fbse_ctx;

/* n_in <= 64, width a multiple of 16 and <= 128, n_out <= 16.  Weights are
 * initialised from seed (uniform +-sqrt(6/(fan_in+fan_out))), biases 0. */
int fbse_mlp_create(int device, int n_in, int width, int n_out, unsigned long long seed, fbse_mlp **out);
void fbse_mlp_destroy(fbse_mlp *m);
const char *fbse_mlp_last_error(const fbse_mlp *m);
int fbse_mlp_set_stream(fbse_mlp *m, void *hip_stream);
int fbse_mlp_input_width(const fbse_mlp *m); /* k_in */
long fbse_mlp_num_params(const fbse_mlp *m);
int fbse_mlp_get_params(fbse_mlp *m, float *params);
int fbse_mlp_set_params(fbse_mlp *m, const float *params);

/* y = net(x) for n points (device pointers). */
int fbse_mlp_forward_dev(fbse_mlp *m, const unsigned short *d_x, long n, float *d_y);

/* Mean-squared error of net(x) against targets t (n x n_out FP32) weighted by
 * w (n x n_out FP32, may be NULL for all ones): loss = sum w (y - t)^2 /
 * sum w.  Leaves the gradient of the loss with respect to every parameter in
 * the module; *loss (may be NULL) receives the loss, grads (may be NULL, host,
 * fbse_mlp_num_params floats) a copy of the gradient in the parameter layout. */
int fbse_mlp_loss_grad_dev(fbse_mlp *m, const unsigned short *d_x, const float *d_t, const float *d_w, long n,
                           double *loss, float *grads);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the gradient left by
 * fbse_mlp_loss_grad_dev. */
int fbse_mlp_adam_step(fbse_mlp *m, float lr);

/* Features of every grid point of a dot into d_x (ngrid x k_in BF16):
 * for each species s < n_species the 8 nearest atoms of that species, as
 * 1/(1 + d/2 bohr) in ascending distance (absent atoms give 0); then V_loc at
 * the point (pot_local of spin 0); then |r - center| / r_scale.  Grid point
 * (jx, jy, jz) sits at origin + dx*(jx, jy, jz) with jx fastest, the
 * reference's order (init.c:592-595).  atoms_xyz (3 per atom) and species are
 * host arrays.  Needs 8*n_species + 2 <= k_in. */
int fbse_mlp_features_dev(fbse_mlp *m, long nx, long ny, long nz, const double *origin, double dx,
                          const double *atoms_xyz, const int *species, long n_atoms, int n_species,
                          const double *d_vloc, const double *center, double r_scale, unsigned short *d_x);

/* Density of n states (reference layout, one state = stlen doubles of the
 * context's grid) summed over spins: d_rho[g] = sum_s sum_sigma |psi_s(g, sigma)|^2
 * over the state indices idx (host array). */
int fbse_state_density_dev(struct fbse_ctx *ctx, const double *d_psi, const long *idx, long n, double *d_rho);

/* Warm start: start state jns of psi_rank (offset jns*ms*stlen, the layout
 * fbse_filter reads) is, when smooth_ep > 0, first low-pass filtered in
 * k-space, F^-1[F[psi] / (1 + 0.5|k|^2 / smooth_ep)] (the refinement
 * preconditioner's FFT pair), which leaves smooth random envelopes on the
 * length scale of the dot; then every grid point and spin is multiplied by
 * exp(0.5 * clamp(y[g*ld + channel], -30, 30)), the square root of the
 * predicted density, and the state is renormalised (sum |psi|^2 dv = 1). */
int fbse_warm_start_dev(struct fbse_ctx *ctx, const float *d_y, int ld, int channel, double smooth_ep,
                        double *d_psi_rank, long n_starts, long ms);

#ifdef __cplusplus
}
#endif

#endif
