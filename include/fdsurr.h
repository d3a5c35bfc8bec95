/* fdsurr.h - C ABI of the gfx950 BF16 MFMA kernels of the convolutional
 * surrogate (3x3 convolutions of a U-Net over the coarse phase field).
 *
 * No reference counterpart: the reference has no learned model.  These are
 * the building blocks of the surrogate that fast-forwards the coarse phase
 * field of the lattice-Boltzmann solver (libfdlbm.so).
 *
 * Conventions
 *   activations  NHWC, bfloat16 (raw uint16 bit patterns), C % 8 == 0
 *   weights      [Cout][3][3][Cin] bfloat16, Cout % 16 == 0
 *   padding      1 (zeros), stride 1: output H x W equals the input's
 *   accumulation FP32 (MFMA v_mfma_f32_16x16x32_bf16)
 *   pointers     device memory owned by the caller; `stream` is a
 *                hipStream_t (NULL: the default stream); every call is
 *                asynchronous on it
 *   returns      0, or a negative code with a message from fds_last_error()
 */
#ifndef FDSURR_H
#define FDSURR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDS_OK 0
#define FDS_E_ARG (-1)
#define FDS_E_HIP (-2)

/* y = act(conv3x3(x, w) + bias): y32 (float NHWC) and/or y16 (bf16 NHWC),
 * either may be NULL; bias may be NULL; relu != 0 applies max(0, .). */
int fds_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const float* bias, float* y32, uint16_t* y16,
                    int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t relu, void* stream);

/* wt[ci][tap][co] = w[co][8 - tap][ci]: the weights of the transposed
 * convolution, so that conv3x3(dy, wt) is the gradient of the input. */
int fds_flip_weights(const uint16_t* w, uint16_t* wt, int32_t Cin, int32_t Cout, void* stream);

/* For n <= 16 layers at once, in one launch: w16[l] = w[l] rounded to bf16
 * (nearest even) and wt[l] = fds_flip_weights of it, from the FP32 weights
 * w[l] [cout[l]][3][3][cin[l]].  w, w16, wt, cin and cout are host arrays of
 * n entries (device pointers in the first three). */
int fds_cast_flip(int32_t n, const float* const* w, uint16_t* const* w16, uint16_t* const* wt, const int32_t* cin,
                  const int32_t* cout, void* stream);

/* dx = conv3x3(dy', wt) with wt from fds_flip_weights (dy has Cout
 * channels, dx has Cin).  dy' = dy where relu_y > 0 and 0 elsewhere when
 * relu_y (the bf16 output of a ReLU forward, same shape as dy) is given,
 * dy otherwise.  dx32 and/or dx16 may be NULL. */
int fds_conv3x3_bwd_data(const uint16_t* dy, const uint16_t* wt, const uint16_t* relu_y, float* dx32,
                         uint16_t* dx16, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, void* stream);

/* Number of floats of scratch fds_conv3x3_bwd_weight needs. */
int64_t fds_bwd_weight_scratch(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout);

/* dw[co][tap][ci] = sum over pixels of dy'[p][co] * x[p + d(tap)][ci]
 * (float, overwritten); dbias[co] = sum over pixels of dy'[p][co] (float,
 * overwritten, may be NULL); dy' as for fds_conv3x3_bwd_data (relu_y may
 * be NULL).  scratch: fds_bwd_weight_scratch floats.  Deterministic:
 * fixed-order partial sums, no atomics. */
int fds_conv3x3_bwd_weight(const uint16_t* x, const uint16_t* dy, const uint16_t* relu_y, float* dw, float* dbias,
                           float* scratch, int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout, void* stream);

/* Decoder input of the U-Net: out[n][h][w][0:C1] = lo[n][h/2][w/2][:]
 * (nearest-neighbour upsampling, lo is N x H/2 x W/2 x C1) and
 * out[n][h][w][C1:C1+C2] = skip[n][h][w][:]; bf16 NHWC, H and W even, C1
 * and C2 multiples of 8.  Replaces torch.cat([up(lo), skip], dim=3). */
int fds_up_cat(const uint16_t* lo, const uint16_t* skip, uint16_t* out, int32_t N, int32_t H, int32_t W,
               int32_t C1, int32_t C2, void* stream);

/* Gradients of fds_up_cat: dlo (N x H/2 x W/2 x C1) = the sum of dout over
 * each 2 x 2 block of its first C1 channels, the two columns of each row
 * summed and rounded to bf16 first, then the two rows (the order of the
 * gradients of two nearest-neighbour repeats, along w then h); dskip = the
 * last C2 channels of dout. */
int fds_up_cat_bwd(const uint16_t* dout, uint16_t* dlo, uint16_t* dskip, int32_t N, int32_t H, int32_t W,
                   int32_t C1, int32_t C2, void* stream);

/* fds_up_cat_bwd with dlo zeroed where lo (the coarse input, the bf16 output
 * of a ReLU, N x H/2 x W/2 x C1) is not > 0: that ReLU's backward.  lo may
 * be NULL (then fds_up_cat_bwd).  dskip may be NULL: it is then not formed,
 * and the caller reads the skip gradient as the channels C1 .. C1 + C2 - 1
 * of dout in place. */
int fds_up_cat_bwd_masked(const uint16_t* dout, const uint16_t* lo, uint16_t* dlo, uint16_t* dskip, int32_t N,
                          int32_t H, int32_t W, int32_t C1, int32_t C2, void* stream);

/* out = dy where y > 0 (bf16 values as floats: -0 and NaN are not), else 0:
 * the backward mask of a ReLU whose bf16 output is y; n values, n % 8 == 0;
 * out may be dy. */
int fds_relu_mask(const uint16_t* dy, const uint16_t* y, uint16_t* out, int64_t n, void* stream);

/* Gradient of s (N x H x W x C bf16, H and W even, C a multiple of 8), the
 * output of a ReLU that feeds a 2 x 2 average pool (gradient dp, N x H/2 x
 * W/2 x C) and a skip connection (gradient dskip, N x H x W x C with pixels
 * dskip_ld channels apart, dskip_ld >= C a multiple of 8, so that a channel
 * slice of the decoder input's gradient is read in place):
 * out = 0 where s is not > 0, else bf16(bf16(dp[n][h/2][w/2][c] / 4) +
 * dskip[n][h][w][c]) -- the mean's backward, the sum of the two gradients
 * and the ReLU's backward, each rounded as torch rounds them. */
int fds_pool_skip_bwd(const uint16_t* dp, const uint16_t* dskip, int32_t dskip_ld, const uint16_t* s, uint16_t* out,
                      int32_t N, int32_t H, int32_t W, int32_t C, void* stream);

/* Message of the last failed call on this thread. */
const char* fds_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FDSURR_H */
