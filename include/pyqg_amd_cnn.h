/* C-ABI of libqgcnn: BF16 MFMA kernels of the learned subgrid closure.
 *
 * No counterpart in the reference: pyqg's parameterizations are Python
 * callbacks (pyqg/parameterizations.py:4-130) plugged into
 * _do_q_subgrid_parameterization (pyqg/kernel.pyx:424-442); this library
 * provides the kernels of a CNN used as such a q parameterization.
 *
 * Tensors are device pointers.  Activations are NHWC BF16 (raw 16-bit
 * patterns), channels padded to a multiple of 16 (<= 128), W = 64, periodic
 * boundaries; weights W[cout][25][cin] BF16 for 5x5 kernels (tap = dy*5+dx);
 * biases and gradients FP32.  `stream` is a hipStream_t (NULL: default).
 * Functions return 0 or a negative QGCNN_E_* code; qgcnn_last_error() explains.
 */
#ifndef PYQG_AMD_CNN_H
#define PYQG_AMD_CNN_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { QGCNN_OK = 0, QGCNN_E_ARG = -1, QGCNN_E_HIP = -2 };

const char* qgcnn_last_error(void);

/* y = act(conv5x5(x, w) + bias), act = ReLU if relu; if mask != NULL the result
 * is zeroed where mask <= 0 (backward through a ReLU).  Writes BF16 y and/or
 * FP32 y32 (either may be NULL).  Every pointer given must be 16-byte aligned
 * (QGCNN_E_ARG otherwise); the same holds for qgcnn_conv_first. */
int qgcnn_conv_fwd(const void* x, int B, int H, int W, int cinp, const void* w,
                   const float* bias, int coutp, int relu, const void* mask, void* y, float* y32,
                   void* stream);
/* First layer of a network with at most 8 real input channels: the same
 * y = act(conv5x5(x, w) + bias) with the 25 taps x 8 channels packed into the
 * matrix K dimension.  x keeps its cinp-channel layout (channels 8.. unused);
 * w8 is w repacked by qgcnn_pack_first: [coutp][28][8] BF16 (device,
 * coutp * 448 bytes), refreshed whenever w changes. */
int qgcnn_pack_first(const void* w, int cinp, int coutp, void* w8, void* stream);
int qgcnn_conv_first(const void* x, int B, int H, int W, int cinp, const void* w8,
                     const float* bias, int coutp, int relu, void* y, float* y32, void* stream);
/* Last layer with cout real output channels, 5 * cout <= 16, cinp a multiple
 * of 32, no activation: y = conv5x5(x, w) + bias over all coutp channels
 * (channels >= cout get their bias), with the 5 horizontal taps of each real
 * output channel as rows of the matrix tile.  wl is w repacked by
 * qgcnn_pack_last: [16][5][cinp] BF16 (device, 160 * cinp bytes). */
int qgcnn_pack_last(const void* w, int cinp, int coutp, int cout, void* wl, void* stream);
/* The output may be narrower than the other layers' (coutp a multiple of 4,
 * at least cout): with coutp = 4 the FP32 output of a 2-channel last layer
 * writes 16 bytes per pixel instead of 64. */
int qgcnn_conv_last(const void* x, int B, int H, int W, int cinp, const void* wl,
                    const float* bias, int coutp, int cout, void* y, float* y32, void* stream);
/* dw[co][tap][ci] += sum_px dy[px][co] x[px + tap][ci] (FP32 accumulation) */
int qgcnn_conv_wgrad(const void* dy, const void* x, int B, int H, int W, int cinp, int coutp,
                     float* dw, void* stream);
/* The same weight gradient with the matrix cores, deterministic: every block
 * writes its partial sums to its own slice of `workspace` (at least
 * qgcnn_wgrad_workspace_bytes bytes, device memory) and a second kernel adds
 * them into dw.  Returns 0 bytes for unsupported shapes. */
size_t qgcnn_wgrad_workspace_bytes(int B, int H, int W, int cinp, int coutp);
int qgcnn_conv_wgrad_ws(const void* dy, const void* x, int B, int H, int W, int cinp, int coutp,
                        float* dw, void* workspace, size_t ws_bytes, void* stream);
/* wt[ci][24 - tap][co] = w[co][tap][ci] (weights of the data-gradient convolution) */
int qgcnn_flip_transpose(const void* w, void* wt, int cinp, int coutp, void* stream);
/* db[c] += sum_px dy[px][c] */
int qgcnn_bias_grad(const void* dy, size_t npix, int coutp, float* db, void* stream);
/* As qgcnn_bias_grad with a fixed summation order (repeated runs agree bit for
 * bit): ws holds at least 256 * coutp floats of per-block partial sums. */
int qgcnn_bias_grad_ws(const void* dy, size_t npix, int coutp, float* db, float* ws,
                       size_t ws_floats, void* stream);
/* loss += mean over px and the first nch channels of (pred - target)^2;
 * grad = d loss / d pred * gscale as BF16 over all coutp channels (may be NULL).
 * pred is [npix][coutp] FP32, target [npix][nch] FP32. */
int qgcnn_mse(const float* pred, const float* target, size_t npix, int coutp, int nch,
              float gscale, void* grad, float* loss, void* stream);
int qgcnn_f32_to_bf16(const float* a, void* o, size_t n, void* stream);
int qgcnn_bf16_to_f32(const void* a, float* o, size_t n, void* stream);
/* Adam (decoupled weight decay off when wd = 0) on FP32 masters p; refreshes the
 * BF16 copy pb when not NULL.  step counts from 1. */
int qgcnn_adam(float* p, const float* g, float* m, float* v, void* pb, size_t n, float lr,
               float b1, float b2, float eps, int step, float wd, void* stream);

#ifdef __cplusplus
}
#endif
#endif
