/*
 * raft_nn.h -- device-side building blocks of the learned eEM warm start: a
 * 4-level 2-D U-Net (3x3 convolutions on BF16 MFMA with FP32 accumulation)
 * that maps a sparse-view FBP image, resampled onto the eEM grid, to a
 * starting image for get_eEM_RT_MultiGPU / raft_eem_device.
 *
 * None of these entry points has a counterpart in the reference (ssc-raft has
 * no learned components); they exist for callers that train or run the warm
 * start network (sscRaft/learned.py).  All pointers are device pointers
 * (hipMalloc'd or torch tensors), all calls are asynchronous on `stream` (a
 * hipStream_t, null = default stream).  Return value: RAFT_OK, RAFT_EINVAL for
 * shapes / alignments the kernels do not accept, RAFT_EHIP for a launch error.
 *
 * Tensors are NHWC bf16 (raw uint16 bit patterns) addressed as
 * base[pixel * ld + c0 + channel], pixel = (n * H + h) * W + w; conv weights
 * are [Cout][9][Cin] with tap = 3 * (dy + 1) + (dx + 1).
 */
#ifndef RAFT_NN_H
#define RAFT_NN_H

#ifdef __cplusplus
extern "C" {
#endif

/* Raw per-lane results of one v_mfma_f32_16x16x32_bf16 (A16 [16][32], B16
 * [32][16] -> out16[64][4]) and one v_mfma_f32_32x32x16_bf16 (A32 [32][16],
 * B32 [16][32] -> out32[64][16]) on row-major bf16 operands. */
int raft_nn_mfma_probe(const unsigned short* A16, const unsigned short* B16, float* out16, const unsigned short* A32,
                       const unsigned short* B32, float* out32, void* stream);

/* y = act(b + conv3x3(x, w)), zero padding; Cin, Cout multiples of 32;
 * ldx, cx0 multiples of 8; ldy, cy0 multiples of 4; relu 0/1; b may be null.
 * variant 0 = automatic; 1..5 select a one-tile-per-block shape (rows x output
 * channels per block: 8x64, 16x32, 4x32, 16x64, 8x32), 6..9 the persistent
 * form that prefetches the next tile while computing (16x32, 8x64, 16x64,
 * 8x32), 10..13 the same four with the tap-column loop not unrolled.  Any
 * H, W >= 1 (partial tiles are masked). */
int raft_nn_conv3x3(const unsigned short* x, int ldx, int cx0, const unsigned short* w, const float* b,
                    unsigned short* y, int ldy, int cy0, int N, int H, int W, int Cin, int Cout, int relu, int variant,
                    void* stream);

/* raft_nn_conv3x3 with its input assembled on the fly (inference path):
 * input channels [0, split) come from x (channel stride ldx, offset cx0) --
 * at H x W, or, when pool != 0, as the 2x2 average of x at 2H x 2W (the values
 * raft_nn_pool2 would write); channels [split, Cin) come from x2 (ldx2, cx2)
 * at H/2 x W/2 with nearest 2x upsampling (the values raft_nn_up2 would write).
 * split is a multiple of 32 in [32, Cin]; split < Cin needs x2 and even H, W.
 * yp (may be null): also write the 2x2 average of y, the values raft_nn_pool2
 * would compute from it, to yp[(n * H/2 + y) * W/2 + x][cp0 + co] (channel
 * stride ldp; H, W even; ldp, cp0 multiples of 4; persistent variants 6..13). */
int raft_nn_conv3x3_ex(const unsigned short* x, int ldx, int cx0, int pool, const unsigned short* x2, int ldx2,
                       int cx2, int split, const unsigned short* w, const float* b, unsigned short* y, int ldy, int cy0,
                       int N, int H, int W, int Cin, int Cout, int relu, int variant, unsigned short* yp, int ldp,
                       int cp0, void* stream);

/* Last 3x3 convolution (Cout = 32) with the 1x1 output head in its epilogue:
 * pred[p] = res[p] + hb[0] + sum_c hw[c] * bf16(act(conv(x)[p][c])), summed over
 * c in order, then max(., floor_value) when clamp.  Equal to raft_nn_conv3x3
 * followed by raft_nn_head without a target, without writing the 32-channel
 * activations.  variant: 0, or a persistent one-tile form (6, 9, 10, 13). */
int raft_nn_conv3x3_head(const unsigned short* x, int ldx, int cx0, const unsigned short* w, const float* b, int N,
                         int H, int W, int Cin, int relu, int variant, const float* hw, const float* hb,
                         const float* res, float* pred, float floor_value, int clamp, void* stream);

/* First two 3x3 convolutions of the network in one pass (32 channels each):
 * y = act(conv3x3(x1; w, b)) with x1 = act0(conv3x3(bf16(img * scale); w0, b0))
 * computed per tile from the FP32 image inside the halo load, never written.
 * w0 is the first layer's [32][9][ldw0] bf16 weights of which input channel 0
 * is used; w the second layer's [32][9][32].  The first layer's 9 exact
 * products are summed by an MFMA (raft_nn_conv_first sums them in tap order,
 * so a bf16 rounding of x1 may differ).  variant: 0 or 6, 9, 10, 13.  yp, ldp,
 * cp0: the 2x2-pooled output as for raft_nn_conv3x3_ex (yp may be null). */
int raft_nn_conv_first2(const float* img, float scale, const unsigned short* w0, int ldw0, const float* b0, int relu0,
                        const unsigned short* w, const float* b, unsigned short* y, int ldy, int cy0, int N, int H,
                        int W, int relu, int variant, unsigned short* yp, int ldp, int cp0, void* stream);

/* First layer from the FP32 image [N][H][W] (one input channel):
 * y[p][cy0 + co] = act(b[co] + sum_tap bf16(img[p + off(tap)] * scale) * w[(co * 9 + tap) * ldw]),
 * zero padding; w is the layer's [Cout][9][ldw] bf16 weights of which input
 * channel 0 is used.  Cout a multiple of 8, at most 64; ldy, cy0 multiples of 8. */
int raft_nn_conv_first(const float* img, float scale, const unsigned short* w, int ldw, const float* b,
                       unsigned short* y, int ldy, int cy0, int N, int H, int W, int Cout, int relu, void* stream);

/* wt[ci][8 - tap][co] = w[co][tap][ci]: backward-data weights, used as
 * raft_nn_conv3x3(g, ..., wt, ...) with Cin = Cout of the forward layer. */
int raft_nn_conv_weight_flip(const unsigned short* w, unsigned short* wt, int Cout, int Cin, void* stream);

/* Floats of workspace raft_nn_conv3x3_wgrad needs (-1: shape rejected). */
long long raft_nn_wgrad_workspace(int N, int H, int W, int Cin, int Cout);

/* dw[co][tap][ci] (+)= sum_p g[p][co] x[p + off(tap)][ci] (FP32),
 * db[co] (+)= sum_p g[p][co] (db may be null); deterministic (fixed-order
 * partial sums in `work`). */
int raft_nn_conv3x3_wgrad(const unsigned short* x, int ldx, int cx0, const unsigned short* g, int ldg, int cg0,
                          float* dw, float* db, float* work, long long work_floats, int N, int H, int W, int Cin,
                          int Cout, int accumulate, void* stream);

/* The elementwise entry points below move 8 channels at a time: C, every
 * channel stride and every channel offset must be multiples of 8. */

/* out = g * (act > 0) over P pixels x C channels (may run in place). */
int raft_nn_relu_mask(const unsigned short* g, int ldg, int cg0, const unsigned short* act, int lda, int ca0,
                      unsigned short* out, int ldo, int co0, long long P, int C, void* stream);

/* 2x2 average pooling to Ho x Wo. */
int raft_nn_pool2(const unsigned short* x, int ldx, int cx0, unsigned short* y, int ldy, int cy0, int N, int Ho,
                  int Wo, int C, void* stream);

/* Gradient of a post-ReLU skip tensor read by a concatenation and a 2x2
 * average pool: out = (gcat + pool_backward(gpool)) * (act > 0); H, W even. */
int raft_nn_skip_grad(const unsigned short* gcat, int ldc, int cc0, const unsigned short* gpool, int ldp, int cp0,
                      const unsigned short* act, int lda, int ca0, unsigned short* out, int ldo, int co0, int N, int H,
                      int W, int C, void* stream);

/* Nearest 2x upsampling from Hl x Wl (backward = 1: sum of the 4 children). */
int raft_nn_up2(const unsigned short* x, int ldx, int cx0, unsigned short* y, int ldy, int cy0, int N, int Hl, int Wl,
                int C, int backward, void* stream);

/* Network input: x[p][0] = bf16(img[p] * scale), channels 1..31 = 0, ld 32. */
int raft_nn_input(const float* img, unsigned short* x, long long P, float scale, void* stream);

/* Workspace sizes of raft_nn_head with a target: floats, and doubles for the loss. */
long long raft_nn_head_workspace(long long P);
long long raft_nn_head_loss_slots(long long P);

/* 1x1 head 32 -> 1 with residual: pred = res + b + sum_c w[c] h[p][c].
 * With target: loss[0] = sum (pred - target)^2 (FP64), gh = d(mean sq.)/dh
 * (bf16), dw, db (+)= their gradients.  Without target: pred clamped below by
 * floor_value when clamp != 0. */
int raft_nn_head(const unsigned short* h, int ldh, int ch0, const float* w, const float* b, const float* res,
                 float* pred, const float* target, unsigned short* gh, int ldgh, int cgh0, float* dw, float* db,
                 double* loss, float* work, long long P, float floor_value, int clamp, int accumulate, void* stream);

/* Adam step t (>= 1) on FP32 parameters with gradient scale gscale; refreshes
 * the bf16 copy pb when non-null. */
int raft_nn_adam(float* p, const float* g, float* m, float* v, unsigned short* pb, long long n, float lr, float b1,
                 float b2, float eps, int t, float gscale, void* stream);

int raft_nn_to_bf16(const float* in, unsigned short* out, long long n, void* stream);

/* Bilinear resampling (edge-clamped) of S images Hi x Wi onto Ho x Wo:
 * out[s][i][j] = img[s](ay + by i, ax + bx j) in input pixel units. */
int raft_nn_resample(const float* img, float* out, int S, int Hi, int Wi, int Ho, int Wo, double ax, double bx,
                     double ay, double by, void* stream);

#ifdef __cplusplus
}
#endif

#endif
