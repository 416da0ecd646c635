/* libotfwi_unet.so - kernels of the learned preconditioner: a 4-level U-Net
 * (32/64/128/256 channels) with BF16 weights and activations and FP32
 * accumulation on the gfx950 matrix cores.
 *
 * The reference has no learned part; none of these entry points replaces a
 * reference interface.  Callers are the inversion driver and its training
 * loop (otfwi_amd/unet.py), which own the device memory.
 *
 * Conventions: tensors are NHWC ([n][y][x][c]) BF16 bit patterns in device
 * memory, 16-byte aligned; convolution weights are [Cout][9][Cin] with tap
 * 3*dy + dx; `stream` is a hipStream_t (NULL = default stream).  Every call
 * only enqueues work and returns 0, or a negative code with a message in
 * otfwi_unet_last_error() when the shapes are not supported.
 */
#ifndef OTFWI_UNET_H
#define OTFWI_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint16_t otfwi_bf16;

const char *otfwi_unet_last_error(void);

/* 3x3 "same" convolution of the channel concatenation [A | B] (B read at
 * (y/2, x/2) when b_half, i.e. nearest-upsampled; CB = 0 for one input):
 * Y = relu?(conv + bias), then zeroed where mask <= 0 (mask may be NULL).
 * CA + CB is 16 or a multiple of 32; Cout a multiple of 32; W a multiple of
 * 32 and H of 256/64 (W % 64 == 0) or 256/32.  With weights from
 * otfwi_unet_flip_weights and A = the output gradient it is the data
 * gradient of a convolution. */
int otfwi_unet_conv3x3(const otfwi_bf16 *A, int CA, const otfwi_bf16 *B, int CB, int b_half,
                       const otfwi_bf16 *Wt, const float *bias, const otfwi_bf16 *mask,
                       otfwi_bf16 *Y, int N, int H, int W, int Cout, int relu, void *stream);

/* Pixel tiles per work-group of otfwi_unet_conv3x3 where the input channels
 * fit one 32- or 16-channel chunk and the output block is 32 channels: 0 (the
 * default) = 8 when the launch keeps >= 1024 work-groups, else 1; 1; or 8.
 * With 8, a work-group stages the weights once and fetches the next two
 * tiles while computing the current one.  Results are bitwise the same for every
 * value.  Process-wide.  No reference counterpart (the reference's network
 * runs in PyTorch). */
int otfwi_unet_set_conv_tiles(int nt);

/* Weight gradient of that convolution: dW[co][tap][ci] += sum over pixels of
 * G[p][co] * input[p + tap][ci], and (dB may be NULL) the bias gradient
 * dB[co] += sum over pixels of G[p][co] (FP32, atomically added: zero both
 * first).  tiles_per_wg = 256-pixel tiles summed per work-group before its
 * atomics. */
int otfwi_unet_wgrad(const otfwi_bf16 *A, int CA, const otfwi_bf16 *B, int CB, int b_half,
                     const otfwi_bf16 *G, float *dW, float *dB, int N, int H, int W, int Cout,
                     int tiles_per_wg, void *stream);

/* The same weight gradient, bitwise reproducible: each work-group stores its
 * partial sums to its row of `work` and a fixed-order reduction (sums of 16
 * rows, repeated) adds them to dW and dB.  work_elems must be at least
 * otfwi_unet_wgrad_work_elems(N, H, W, CA + CB, Cout, tiles_per_wg) floats;
 * work = NULL falls back to the atomics of otfwi_unet_wgrad.  Calls sharing a
 * workspace must be ordered (one stream).  No reference counterpart (the
 * reference's network runs in PyTorch). */
int otfwi_unet_wgrad_work(const otfwi_bf16 *A, int CA, const otfwi_bf16 *B, int CB, int b_half,
                          const otfwi_bf16 *G, float *dW, float *dB, int N, int H, int W, int Cout,
                          int tiles_per_wg, float *work, size_t work_elems, void *stream);
/* Workspace floats otfwi_unet_wgrad_work needs for this shape (-1: bad shape). */
long long otfwi_unet_wgrad_work_elems(int N, int H, int W, int Cin, int Cout, int tiles_per_wg);

/* Wf[ci][tap][co] = W[co][8 - tap][ci]. */
int otfwi_unet_flip_weights(const otfwi_bf16 *W, int Cout, int Cin, otfwi_bf16 *Wf, void *stream);

/* 2x2 average pooling (C a multiple of 8). */
int otfwi_unet_pool2(const otfwi_bf16 *X, int N, int H, int W, int C, otfwi_bf16 *Y, void *stream);

/* Backward of nearest upsampling then ReLU: out (Ho x Wo x C) = mask > 0 ?
 * sum of the 2x2 fine pixels of src channels [off, off + C) : 0. */
int otfwi_unet_upsum_mask(const otfwi_bf16 *src, int Cs, int off, const otfwi_bf16 *mask, int N,
                          int Ho, int Wo, int C, otfwi_bf16 *out, void *stream);

/* Gradient of an encoder output: out = mask > 0 ? src[off + c] +
 * below[y/2][x/2][c] / 4 : 0. */
int otfwi_unet_skip_combine(const otfwi_bf16 *src, int Cs, int off, const otfwi_bf16 *below,
                            const otfwi_bf16 *mask, int N, int H, int W, int C, otfwi_bf16 *out,
                            void *stream);

/* (N, 2, H, W) FP32 planes -> (N, H, W, 16) BF16, channels 2..15 zero. */
int otfwi_unet_pack_input(const float *in, int N, int H, int W, otfwi_bf16 *x0, void *stream);

/* Output head: out[p] = b + sum_c X[p][c] * w[c], C = 32, FP32 out. */
int otfwi_unet_head_fwd(const otfwi_bf16 *X, const float *w, const float *b, long long np,
                        float *out, void *stream);

/* Backward of the head under the loss sum (out - target)^2 * scale/2:
 * g = scale*(out - target); gX = X > 0 ? g*w : 0; dw += sum g*X; db += sum g;
 * loss += sum (out - target)^2 (FP64). */
int otfwi_unet_head_bwd(const otfwi_bf16 *X, const float *w, const float *out,
                        const float *target, long long np, float scale, otfwi_bf16 *gX, float *dw,
                        float *db, double *loss, void *stream);

/* The same with the block sums stored to `work` (at least
 * OTFWI_UNET_HEAD_WORK floats) and added in block order: bitwise
 * reproducible.  work = NULL: the atomics of otfwi_unet_head_bwd.  No
 * reference counterpart. */
#define OTFWI_UNET_HEAD_WORK (1024 * 34)
int otfwi_unet_head_bwd_work(const otfwi_bf16 *X, const float *w, const float *out,
                             const float *target, long long np, float scale, otfwi_bf16 *gX,
                             float *dw, float *db, double *loss, float *work, size_t work_elems,
                             void *stream);

/* dB[c] += sum over np pixels of G[p][c] (C divides 256). */
int otfwi_unet_bias_grad(const otfwi_bf16 *G, long long np, int C, float *dB, void *stream);

/* Adam step (bias-corrected, step >= 1) on FP32 parameters; pb (may be
 * NULL) receives the BF16 copy. */
int otfwi_unet_adam(float *p, const float *g, float *m, float *v, otfwi_bf16 *pb, long long n,
                    float lr, float b1, float b2, float eps, int step, void *stream);

/* y = BF16(x). */
int otfwi_unet_to_bf16(const float *x, long long n, otfwi_bf16 *y, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OTFWI_UNET_H */
