/*
 * lbm3d_conv.h -- C ABI of the 3-D convolution kernels of liblbm3d.so: BF16
 * implicit-GEMM convolutions on the gfx950 matrix cores (MFMA) with FP32
 * accumulation, forward and backward, for the warm-start surrogate.
 *
 * No reference counterpart: the reference initialises every run at rest
 * (init, Single_phase/LBM_3D_SinglePhase_Solver.py:160-170).  The surrogate
 * predicts steady fields from the geometry and lbm3d_init starts the solver
 * at their equilibrium instead.
 *
 * Layouts (all device pointers, caller-owned; nothing is allocated):
 *   activations  BF16, channels last: x[((b*D + z)*H + y)*W + x][c], for a batch
 *                of `n` volumes of D x H x W voxels and C channels;
 *   weights      BF16 w[co][k], k = tap*cin + ci, tap = (dz+1)*9 + (dy+1)*3 +
 *                (dx+1) for ks = 3 (tap 0 for ks = 1), rows padded with zeros
 *                to kpad = lbm3d_conv_kpad(cin, ks) (a multiple of 32);
 *   bias         FP32 b[cout] (may be NULL);
 *   weight grads FP32 dw[co][kpad].
 * Convolutions are "same" (stride 1, zero padding ks/2, or periodic wrap for
 * the forward convolution with LBM3D_CONV_PERIODIC).  cin and cout must
 * be multiples of 16, at most 256; a batch holds fewer than 2^31 voxels.
 * Return 0 or a negative LBM3D_E* code (lbm3d.h); lbm3d_conv_last_error()
 * gives the message.
 */
#ifndef LBM3D_CONV_H
#define LBM3D_CONV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *lbm3d_conv_last_error(void);

/* Padded row length of the weight matrix: ks^3 * cin rounded up to 32. */
int32_t lbm3d_conv_kpad(int32_t cin, int32_t ks);

/* Flags of lbm3d_conv3d_fwd. */
#define LBM3D_CONV_RELU 1     /* apply ReLU to the output */
#define LBM3D_CONV_PERIODIC 2 /* wrap taps past a face to the opposite face instead of zero padding */

/* y = act(conv(x, w) + bias), act = ReLU with LBM3D_CONV_RELU, BF16 out (FP32
 * accumulation); with LBM3D_CONV_PERIODIC each volume is treated as one period
 * of a periodic field (no padding needed around periodic geometries). */
int lbm3d_conv3d_fwd(void *stream, const uint16_t *x, const uint16_t *w, const float *bias, uint16_t *y,
                     int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t ks,
                     int32_t flags);

/* Weights of the data-gradient convolution: wt[ci][tap'*cout + co] =
 * w[co][tap*cin + ci] with tap' = ks^3 - 1 - tap (the kernel flipped and
 * its channel roles swapped), rows padded to lbm3d_conv_kpad(cout, ks). */
int lbm3d_conv3d_wflip(void *stream, const uint16_t *w, uint16_t *wt, int32_t cin, int32_t cout, int32_t ks);

/* dx = conv(dy, wt): the gradient of a convolution's input, given its output
 * gradient dy (BF16, cout channels) and the flipped weights of lbm3d_conv3d_wflip. */
int lbm3d_conv3d_dgrad(void *stream, const uint16_t *dy, const uint16_t *wt, uint16_t *dx, int32_t n,
                       int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t ks);

/* dw[co][tap*cin + ci] = sum over voxels of dy[v][co] * x[v + off(tap)][ci]
 * (FP32; the padding columns are set to 0), through per-block partial sums in
 * `workspace` of lbm3d_conv3d_wgrad_workspace_bytes() bytes, folded in a fixed
 * order (deterministic). */
size_t lbm3d_conv3d_wgrad_workspace_bytes(int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cin,
                                          int32_t cout, int32_t ks);
int lbm3d_conv3d_wgrad(void *stream, const uint16_t *x, const uint16_t *dy, float *dw, int32_t n, int32_t d,
                       int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t ks, void *workspace,
                       size_t workspace_bytes);

/* Backward of a fused bias + ReLU: dz = g * (y > 0) elementwise (y the
 * layer's BF16 output; a masked element is g * 0.0 as PyTorch computes it),
 * or dz = g when y is NULL, and db[c] = sum over voxels of dz[v][c] in FP32,
 * through per-block partial sums in `workspace` folded in a fixed order.
 * g, y, dz: [voxels][c] BF16 (dz may be g); c divides 2048, 8..256. */
size_t lbm3d_relu_bias_grad_workspace_bytes(int64_t voxels, int32_t c);
int lbm3d_relu_bias_grad(void *stream, const uint16_t *g, const uint16_t *y, uint16_t *dz, float *db,
                         int64_t voxels, int32_t c, void *workspace, size_t workspace_bytes);

/* The surrogate's distance input: dist[v] = Euclidean distance from voxel v to
 * the nearest solid voxel (solid[v] != 0) of its volume, with every volume
 * periodic, exact up to `margin` voxels and capped there (also when a volume
 * has no solid).  n volumes of d x h x w voxels, C order; `work` is a float
 * buffer of the same size.  Three separable min-plus passes over offsets
 * |o| <= margin; margin 1..1024. */
int lbm3d_distance_periodic(void *stream, const int8_t *solid, float *dist, float *work, int32_t n, int32_t d,
                            int32_t h, int32_t w, int32_t margin);

#ifdef __cplusplus
}
#endif

#endif /* LBM3D_CONV_H */
