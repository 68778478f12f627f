/* Warm-start network for the MI355X White-Bear minimizer.
 *
 * A periodic 3-D U-Net (four levels, 3x3x3 convolutions, 2x2x2 average
 * pooling, nearest upsampling, additive skips, ReLU, 1x1 head) maps the
 * Boltzmann factor exp(-Vext/kT) of a cubic M^3 grid and the bulk packing
 * fraction eta to u = ln(n / n_bulk), an estimate of the equilibrium density.
 * The convolutions are BF16 MFMA implicit GEMMs with FP32 accumulation and
 * FP32 activations; training (masked mean squared error on u, Adam) runs on the
 * device with the library's own backward kernels.
 *
 * None of these entry points has a counterpart in deft: deft starts every
 * minimization from a bulk or hand-built Veff (e.g. papers/contact/figs/
 * inner-sphere.cpp sets Veff = -kT ln(n_bulk) everywhere) and runs conjugate
 * gradients from there.  deft_plan_warm_start replaces that initial Veff;
 * the conjugate-gradient minimizer still decides convergence.
 *
 * Device buffers (_dev) are device pointers from deft_buffer_create (or any
 * hipMalloc of the same HIP runtime); the library works on its own stream and
 * synchronises it before returning, so the caller must have finished writing
 * the inputs.
 */
#ifndef DEFT_AMD_NN_H
#define DEFT_AMD_NN_H

#include "deft_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct deft_net_s *deft_net;

/* channels: four multiples of 16 in [16, 64] (NULL = 16, 32, 48, 64);
 * weights are He-uniform from a fixed-seed generator. */
int deft_net_create(const int channels[4], unsigned long long seed, deft_net *out);
int deft_net_destroy(deft_net net);

/* tiled = 1 (default): LDS-tiled convolutions on levels whose edge is a
 * multiple of 16, the untiled kernel elsewhere; tiled = 0: the untiled kernel
 * everywhere.  Both compute identical sums in the same order. */
int deft_net_set_conv_kernel(deft_net net, int tiled);

/* number of FP32 parameters, and where convolution `conv` (0..13, or -1 for
 * the head) keeps them: weights W[co][k] at woff with row length kp
 * (k = tap * cin + ci, tap = 9 (dx+1) + 3 (dy+1) + (dz+1), zero padding past
 * 27 cin), biases at boff. */
long long deft_net_num_params(deft_net net);
int deft_net_layout(deft_net net, int conv, int *cin, int *cout, int *kp, long long *woff, long long *boff);
int deft_net_get_params(deft_net net, float *out_host);
/* replaces the parameters and clears the Adam moments */
int deft_net_set_params(deft_net net, const float *in_host);

/* u_dev[b][x][y][z] for a batch of `batch` M^3 grids; boltz_dev holds
 * exp(-Vext/kT) in the same layout, eta_host one packing fraction per grid.
 * M is a multiple of 8 in [16, 512]. */
int deft_net_forward(deft_net net, int M, int batch, const float *boltz_dev, const float *eta_host, float *u_dev);

/* loss = sum over voxels with mask > 0 of (u - target)^2 / count (mask_dev
 * NULL: the voxels with boltz > 0); the
 * gradient with respect to every parameter is left in the network and, when
 * grads_host is not NULL, copied there (same layout as the parameters). */
int deft_net_loss_grad(deft_net net, int M, int batch, const float *boltz_dev, const float *eta_host,
                       const float *target_dev, const float *mask_dev, float *loss, float *grads_host);

/* loss_grad followed by one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) */
int deft_net_train_step(deft_net net, int M, int batch, const float *boltz_dev, const float *eta_host,
                        const float *target_dev, const float *mask_dev, float lr, float *loss);

/* Device buffers for callers without their own HIP allocator. */
int deft_buffer_create(unsigned long long bytes, void **dev);
int deft_buffer_destroy(void *dev);
int deft_buffer_upload(void *dev, unsigned long long offset, const void *host, unsigned long long bytes);
int deft_buffer_download(const void *dev, unsigned long long offset, void *host, unsigned long long bytes);

/* Training-batch gatherer: out_dev[b] = data_dev[samples[b]] (count M^3 grids)
 * under a cubic symmetry and periodic shift; transforms holds 9 ints per grid:
 * an axis permutation p[3], flips f[3] (0/1) and shifts s[3] in [0, M).  Output
 * axis a reads input axis p[a] at f[a] ? M-1-((o-s[a]) mod M) : (o-s[a]) mod M.
 * Apply the same transforms to the Boltzmann factors and the targets. */
int deft_net_make_batch(const float *data_dev, long long count, int M, int batch, const int *samples,
                        const int *transforms, float *out_dev);

/* Sets the plan's Veff from the network's prediction for its external
 * potential: Veff = -kT (ln n_bulk + u), n_bulk = eta / (4 pi R^3 / 3).
 * The plan must be cubic with an edge that is a multiple of 8; points with
 * Vext = +inf keep density zero whatever Veff holds.  One energy evaluation
 * checks the start: where the prediction makes it infinite or NaN (a weighted
 * packing fraction at or past 1), u is scaled by lambda = 1/2, 1/4, ... and
 * finally 0 (the bulk start) until the energy is finite. */
int deft_plan_warm_start(deft_plan plan, deft_net net, double eta);
/* The same for a grid finer than the network's training spacing: the network
 * runs on the (edge/factor)^3 grid (the Boltzmann factor smoothed onto it with
 * weights 1/factor per axis over the points within factor/2, half weight at
 * exactly factor/2), and u is interpolated back trilinearly (periodic).
 * factor 1, 2 or 4; factor 1 is deft_plan_warm_start. */
int deft_plan_warm_start_scaled(deft_plan plan, deft_net net, double eta, int factor);
/* lambda of the last warm start (1 = the network's prediction unchanged). */
int deft_net_last_blend(deft_net net, double *lambda);

/* deft_plan_warm_start_scaled for `count` plans of one cubic grid edge at once:
 * their Boltzmann factors form one batch of a single inference (no counterpart
 * in the reference; for batches of minimizations, where one inference per
 * plan would cost more host-driven launches than the start saves).  Each
 * plan's start is checked by its own energy evaluation as in the single-plan
 * call; lambdas (may be null) receives each plan's blend factor. */
int deft_plans_warm_start(const deft_plan *plans, const double *etas, int count, deft_net net, int factor,
                          double *lambdas);

/* One layer on its own, for checking each kernel against a reference (host
 * buffers, channels-last, M^3 grids, batch of them):
 *   op 0: out = act(conv(in) + bias), in B M^3 x cin, w cout x kp (FP32,
 *         rounded to BF16 on the device), out B M^3 x cout; relu 0/1
 *   op 1: backward-data of that convolution: in = d(pre-activation) B M^3 x cout,
 *         out B M^3 x cin (no bias, no activation)
 *   op 2: weight gradient: in = x (B M^3 x cin), in2 = d(pre-activation)
 *         (B M^3 x cout), out cout x kp
 *   op 3: 2x2x2 average pool of in (B M^3 x cin) -> out (B (M/2)^3 x cin)
 *   op 4: nearest upsample of in (B (M/2)^3 x cin) -> out (B M^3 x cin)
 * kp = 27 cin rounded up to 32; tiled selects the LDS-tiled convolution where
 * M is a multiple of 16. */
int deft_net_layer_probe(int op, int M, int batch, int cin, int cout, const float *in_host, const float *in2_host,
                         const float *w_host, const float *bias_host, int relu, int tiled, float *out_host);

/* One 16x16 output tile of the convolution kernel's BF16 MFMA GEMM with
 * K = 32: c[m][n] = sum_k a[m][k] b[n][k] (inputs rounded to BF16). */
int deft_net_mfma_probe(const float *a_host, const float *b_host, float *c_host);

#ifdef __cplusplus
}
#endif

#endif
