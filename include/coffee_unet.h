/* C ABI of the learned warm start: the building blocks of a 2D convolutional
 * U-Net (BF16 activations and weights, FP32 accumulation on MFMA), trained and
 * run on the GPU by the project's own kernels.
 *
 * No entry point here has a counterpart in the reference (fizban007/CoffeeGPU
 * has no learned component); they serve the batched log-spherical solver
 * (coffee_sph_batch_*), whose steady states are the training data and whose
 * runs the trained network warm-starts.
 *
 * Conventions: activations are NHWC with the channel count padded to a
 * multiple of 16 (padded channels hold zeros); BF16 values are stored as
 * uint16 (round-to-nearest-even from FP32); a 3x3 convolution (zero padding 1)
 * has BF16 weights W[tap][cout][cin] (tap = 3*ky + kx) and, for the backward
 * data pass, their transpose WT[tap][cin][cout].  All calls are asynchronous
 * on `stream` unless stated; every pointer is device memory.  Return 0 or a
 * negative COFFEE_ERR_* status; the message is in coffee_last_error(NULL).
 */
#ifndef COFFEE_UNET_H
#define COFFEE_UNET_H

#ifdef __cplusplus
extern "C" {
#endif

/* out[n,y,x,co] = act(bias[co] + sum_{tap,ci} W[tap][co][ci] in[n,y+s*dy,x+s*dx,ci])
 * with (dy, dx) = (ky-1, kx-1), s = +1 (flip = 0, forward) or -1 (flip = 1,
 * backward data pass with WT).  act = ReLU if relu.  out (BF16) and/or out32
 * (FP32) may be NULL; bias may be NULL.  cin, cout multiples of 16. */
int coffee_unet_conv3x3(const void *in, const void *w, const float *bias, void *out, float *out32,
                        int n, int h, int wd, int cin, int cout, int flip, int relu, void *stream);

/* Weight gradient: dW[tap][co][ci] = sum_p dz[p][co] * x[p + (dy,dx)][ci],
 * split over `nsplit` pixel ranges into `partial` (nsplit*9*cout*cin FP32),
 * then reduced into dW (FP32); db[co] = sum_p dz[p][co] (FP32, may be NULL). */
int coffee_unet_conv3x3_wgrad(const void *x, const void *dz, float *partial, int nsplit, float *dW,
                              float *db, int n, int h, int wd, int cin, int cout, void *stream);

/* out[p][ooff+c] = (a[p][aoff+c] + b[p][boff+c]) * (mask[p][moff+c] > 0), c < nc,
 * over p < npix; b and mask may be NULL (then 0 and 1).  *st = channel
 * strides of out, a, b, mask.  Concatenation, split, skip sums and the ReLU
 * backward pass are all this one kernel. */
int coffee_unet_combine(void *out, int ost, int ooff, const void *a, int ast, int aoff,
                        const void *b, int bst, int boff, const void *mask, int mst, int moff,
                        int nc, long npix, void *stream);

/* 2x2 average pooling and its adjoint; nearest 2x upsampling and its adjoint
 * (h, wd are the sizes of the larger grid; c = channel count). */
int coffee_unet_pool2(const void *in, void *out, int n, int h, int wd, int c, void *stream);
int coffee_unet_pool2_bwd(const void *dy, void *dx, int n, int h, int wd, int c, void *stream);
int coffee_unet_up2(const void *in, void *out, int n, int h, int wd, int c, void *stream);
int coffee_unet_up2_bwd(const void *dy, void *dx, int n, int h, int wd, int c, void *stream);

/* Mean squared error of out32 (FP32, channel stride cst) against target (FP32,
 * nc channels, stride nc) over npix pixels; writes the loss gradient
 * 2 (out - target) / (npix*nc) as BF16 into dout (stride cst, padded channels
 * zero) and the loss into *loss (host, synchronous). */
int coffee_unet_mse(const float *out32, int cst, const float *target, int nc, long npix,
                    void *dout, double *loss, void *stream);

/* Adam step on one tensor of n FP32 parameters (t = step number, 1-based);
 * when wb is given the tensor is a 3x3 weight [9][cout][cin] and its BF16
 * copies W and WT are refreshed. */
int coffee_unet_adam(float *p, const float *g, float *m, float *v, long n, float lr, float beta1,
                     float beta2, float eps, int t, void *wb, void *wbt, int cout, int cin,
                     void *stream);

/* FP32 -> BF16 weight copies W[tap][co][ci] and WT[tap][ci][co]. */
int coffee_unet_weights_bf16(const float *w32, void *wb, void *wbt, int cout, int cin,
                             void *stream);

/* FP32 <-> BF16 conversion of n values. */
int coffee_unet_to_bf16(const float *in, void *out, long n, void *stream);
int coffee_unet_from_bf16(const void *in, float *out, long n, void *stream);

/* Log-spherical runs <-> network tensors.  pack: for run b, the 6 field
 * arrays (Ex..Bz, Scalar = scalar_bytes, Fortran order dims0 x dims1 with
 * guards g0, g1) become orthonormal components times r^rpow at the n0 x n1
 * interior nodes, written as channels 0..5 of an NHWC tensor (BF16, or FP32
 * when out_fp32) with channel stride cst, image `image`, H = n1 (theta), W = n0
 * (ln r); channel 6 = extra_channel when cst > 6; other channels are zeroed.  unpack: FP32 NHWC (stride cst, channels 0..5) back into the field
 * arrays' interior, dividing by r^rpow: F^i = value / (r^rpow * h_i) with
 * h = (r, r, r sin(theta)).  r = exp(x0 + dx (i - g0)), theta = dy (j - g1). */
int coffee_unet_pack(void *const fields[6], int scalar_bytes, int dims0, int g0, int g1, int n0,
                     int n1, double x0, double dx, double dy, double rpow, double extra_channel,
                     void *out, int out_fp32, int cst, int image, void *stream);
int coffee_unet_unpack(const float *in, int cst, int image, void *const fields[6],
                       int scalar_bytes, int dims0, int g0, int g1, int n0, int n1, double x0,
                       double dx, double dy, double rpow, void *stream);
/* Corotating electric field from B on a run's interior: E = -(Omega x r) x B,
 * i.e. E^x = v B^theta, E^theta = -v B^x, E^phi = 0 with v = omega r sin(theta)
 * (the relation boundary_pulsar_sph imposes inside the star, and which the
 * steady aligned rotator satisfies on every field line).  Geometry as in
 * coffee_unet_unpack; fields[5] is not read. */
int coffee_unet_corotate(void *const fields[6], int scalar_bytes, int dims0, int g0, int g1,
                         int n0, int n1, double x0, double dx, double dy, double omega,
                         void *stream);

#ifdef __cplusplus
}
#endif

#endif
