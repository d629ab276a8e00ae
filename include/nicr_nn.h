/* nicr_nn.h -- learned surrogate of the step (part of libnicr.so).
 *
 * No counterpart in the reference, which has only the explicit step
 * (mpi2d/NiCr2D_grains.cpp:551-597, mpi3d-slab/NiCr3D_grains.cpp:761-826). A 3D
 * U-Net (4 levels, 32/64/128/256 channels, 3x3x3 convolutions with ReLU, 2x2x2
 * average pooling, nearest upsampling with skip concatenation, FP32 1x1 head;
 * X zero padded, Y/Z periodic like the mesh) maps the state (O, Cr, phi, gb
 * mask) to its change over `jump` steps. Weights and activations are BF16 with
 * FP32 accumulation on MFMA; master weights, gradients and Adam moments are FP32.
 *
 * Conventions as nicr.h: plain C, host arrays borrowed for one call, every
 * function returns NICR_OK or a negative NICR_E_* code (nicr_last_error()).
 * Host volumes are float32 in the model's cell order ((x*ny + y)*nz + z), one
 * channel after another: inputs [nb][7][V] (O, Cr, phi, gb and the plain
 * one-step update of O, Cr, phi), outputs and targets [nb][3][V] (normalised
 * corrections, see nicr_unet_norm).
 */
#ifndef NICR_NN_H
#define NICR_NN_H

#include <stddef.h>
#include <stdint.h>
#include "nicr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A network for volumes of s0 x s1 x s2 cells (multiples of 8), batches of up
 * to max_batch volumes. train != 0 allocates gradients and Adam moments;
 * max_samples > 0 reserves a training set of that many samples on the device.
 * Weights: He-normal from `seed`, zero biases, zero head. */
int nicr_unet_create(int s0, int s1, int s2, int max_batch, int train, int max_samples, uint64_t seed, void** net);
int nicr_unet_destroy(void* net);
int nicr_unet_param_count(void* net, long long* count);
int nicr_unet_device_bytes(void* net, size_t* bytes);
/* FP32 parameters: per convolution W[27][cout][cin] then b[cout] (tap t = 9a +
 * 3b + c has offsets a-1, b-1, c-1 along X, Y, Z) in the order e1a e1b e2a e2b
 * e3a e3b b4a b4b d3a d3b d2a d2b d1a d1b, then the head W[3][32], b[3].
 * set_params also resets the Adam state. */
int nicr_unet_get_params(void* net, float* host);
int nicr_unet_set_params(void* net, const float* host);
/* Gradients of the last nicr_unet_loss_backward (same layout as the parameters). */
int nicr_unet_get_grads(void* net, float* host);
/* Normalisation. Each volume is seen relative to its own statistics (mean m
 * and population standard deviation s of O and Cr, RMS r of the plain one-step
 * update u of O, Cr, phi; binary64). Input channels: 0, 1 = (O - m_O) / s_O,
 * (Cr - m_Cr) / s_Cr; 2, 3 = (phi - phi_mean) / phi_scale, (gb - gb_mean) /
 * gb_scale; 4..6 = u_O / r_O, u_Cr / r_Cr, u_phi / r_phi (0 where the
 * divisor is 0); 7..10 constant fields (c_k - c_mean[k]) / c_scale[k] of c =
 * (log10 s_O, m_Cr, log10 r_O, log10 r_Cr); channels 11..31 are zero. The
 * network predicts the correction to the linear extrapolation of the update
 * over `jump` steps: change of field k = jump u_k + out_k jump r_k gain[k].
 * stat_planes > 0 takes m, s and r over the first stat_planes X planes only
 * (the inlet side), so that a mesh deeper than the training volumes presents
 * the statistics of a training volume's depth (an inlet front fills a smaller
 * share of a deeper mesh); 0 or >= nx: the whole volume. nicr_sur_fit_norm
 * sets it to the training volumes' s0. */
typedef struct nicr_unet_norm {
    float phi_mean, phi_scale, gb_mean, gb_scale;
    float c_mean[4], c_scale[4];
    float gain[3];
    int32_t jump;
    int32_t stat_planes;
} nicr_unet_norm;
int nicr_unet_set_norm(void* net, const nicr_unet_norm* norm);
int nicr_unet_get_norm(void* net, nicr_unet_norm* norm);
/* Forward pass of nb volumes: in [nb][7][V] -> out [nb][3][V] (normalised). */
int nicr_unet_forward(void* net, int nb, const float* in, float* out);
/* Forward pass, then the gradients of L = (1/3) sum_p sum_c (out - target)^2
 * with respect to every parameter (overwritten, not accumulated); *loss = L
 * divided by the number of positions. */
int nicr_unet_loss_backward(void* net, int nb, const float* in, const float* target, double* loss);
/* Device time of the network's forward passes inside nicr_sur_advance. */
int nicr_unet_set_timing(void* net, int enabled);
int nicr_unet_forward_ms(void* net, double* ms);
/* One 3x3x3 convolution of the network's kernels on host arrays, for kernel
 * checks: op 0 forward y = act(conv(x, w) + b) -> out [P][cout]; op 1 data
 * gradient conv^T(dy, w) -> out [P][cin]; op 2 weight gradient -> out
 * [27][cout][cin]. x [P][cin], dy [P][cout] channels-last over P = nb*s0*s1*s2
 * positions (rounded to BF16 on upload), w [27][cout][cin], b [cout];
 * cin, cout multiples of 32. */
int nicr_nn_conv_check(int op, int nb, int s0, int s1, int s2, int cin, int cout, int relu, const float* x,
                       const float* w, const float* b, const float* dy, float* out);

/* ---- training set and training ---- */
/* Record one training sample from a binary64 3D model of the network's volume.
 * The call takes 1 + jump steps on the model (nicr_step): the first gives the
 * plain update u; the sample holds the state after it (O, Cr, phi, mask), u,
 * and the change over the next `jump` steps. All samples of a set share
 * `jump`. A rejected step records nothing and is reported in *report. */
int nicr_sur_record(void* net, void* model, int jump, nicr_step_report* report);
int nicr_sur_samples(void* net, int* count);
/* Normalisation from the training set: phi and gb means and standard
 * deviations, means and standard deviations of the volume statistics, gains
 * that give the normalised corrections unit mean square, and the jump. */
int nicr_sur_fit_norm(void* net);
/* Training samples through the symmetries of the step (1) or as recorded (0,
 * the default): each drawn sample is shifted periodically in Y and Z by random
 * offsets and mirrored in Y and in Z with probability 1/2 each; X, which holds
 * the inlet, is kept. No counterpart in the reference. */
int nicr_sur_set_augment(void* net, int on);
/* `iters` Adam iterations (learning rate lr) on random batches of `batch`
 * samples drawn with `seed`; the gradient is clipped to 3 times the running
 * mean of its global norm, and an iteration with a non-finite gradient leaves
 * the parameters unchanged. losses[iters] (mean squared error of the
 * normalised changes per iteration) when not NULL. */
int nicr_sur_train(void* net, int iters, int batch, float lr, uint64_t seed, float* losses);
/* Mean squared error of the normalised changes over samples [first, first+count). */
int nicr_sur_eval(void* net, int first, int count, double* mse);

/* ---- surrogate-accelerated stepping ---- */
typedef struct nicr_sur_options {
    int32_t jump;             /* steps replaced by one network proposal (50) */
    int32_t conserve_mass;    /* O and Cr totals after a jump follow its sources (see nicr_sur_advance) */
    double residual_tol;      /* accept when ||R|| <= residual_tol * ||proposal - x0|| for O and Cr */
    double min_value, max_value;   /* accept only O, Cr, phi inside [min, max] */
    int32_t settle_steps;     /* extra plain steps before every proposal (0: none) */
    int32_t fallback_steps;   /* plain steps after a rejected jump before the next proposal (0: jump) */
    int32_t min_jump;         /* shortest jump at the end of a call (0: full jumps only; see nicr_sur_advance) */
} nicr_sur_options;

typedef struct nicr_sur_report {
    int32_t steps_done;       /* model steps advanced (plain + jumps x jump) */
    int32_t plain_steps;      /* steps taken by the binary64 step itself */
    int32_t jumps_accepted, jumps_rejected;
    int32_t status;           /* NICR_STEP_* of a rejected plain step, else OK */
    int32_t reserved;
    double max_residual_o, max_residual_cr;   /* largest accepted ||R|| / ||proposal - x0|| */
    double min_rejected_residual;             /* smallest max(O, Cr, eta) residual among rejections */
    double max_residual_eta;                  /* largest accepted eta residual (evolving grains) */
    double last_residual[3];                  /* O, Cr, eta residuals of the last jump attempt (inf: not
                                                 computed, the proposal failed the bounds or the check step) */
    nicr_step_report step;    /* the plain step's report when status != OK */
} nicr_sur_report;

/* Advance a binary64 3D model by nsteps with network jumps of `jump` steps.
 * Each cycle starts from two consecutive states x_-1 -> x0 (update u0 =
 * x0 - x_-1): the first cycle takes 1 + settle_steps plain steps for them;
 * later cycles use the last step taken (the check step of an accepted jump or
 * the last plain step of a rejected one) after settle_steps more plain steps;
 * the proposal
 * xh = x0 + jump u0 + correction(x0, u0) of O, Cr, phi for the state `jump`
 * steps later (jump must equal the network's); evolving eta fields jump to
 * x0 + jump u0 with their mask rebuilt, static ones are unchanged; optionally
 * O and Cr rescaled to their totals (below); one plain check step from xh to
 * xh1. The jump is accepted when O, Cr, phi (and evolving eta) are inside
 * [min, max], the model accepts the check step, and for O, Cr (and all
 * evolving eta fields together) the trapezoid residual
 * R = (xh - x0) - (jump / 2) (u0 + (xh1 - xh)) satisfies
 * ||R|| <= residual_tol ||xh - x0|| (an O or Cr proposal that does not move
 * fails). Otherwise the model is put back exactly to x0 (fields, eta, mask,
 * state index, budgets) and takes fallback_steps (default `jump`) plain steps
 * instead.
 * Sources over a jump: the budget terms of the `jump` steps it replaces (O
 * inlet of the concentration or mu boundary, legacy Ck, clipping) are the
 * trapezoid jump/2 (s_pre + s_check) of the per-step sources of the step that
 * gave u0 and of the check step; they are added to the budgets and the clipped
 * count. With conserve_mass, O and Cr are rescaled to their totals at x0 plus
 * jump s_pre before the check step, and after it by the trapezoid's
 * difference, so the totals follow the budgets (a closed box: the totals at
 * x0). Ksp scans are not run (as nicr_step runs none): the caller ends each
 * call at the next scan and calls nicr_nucleate there.
 * Near the end of a call, fewer steps than a full jump (plus its check step)
 * may be left: with min_jump in [2, jump) a last jump of Jc >= min_jump steps
 * takes them, with the linear extrapolation Jc u0 and the network's
 * correction scaled by (Jc / jump)^2 (the deviation from a linear path is
 * second order in the jump length), checked like any other; otherwise, and
 * when it is rejected, plain steps take them. */
int nicr_sur_advance(void* net, void* model, int nsteps, const nicr_sur_options* opt, nicr_sur_report* report);

#ifdef __cplusplus
}
#endif
#endif /* NICR_NN_H */
