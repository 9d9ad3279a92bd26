/* Internal interface of the learned demosaic (prr_net.hip): tensor
 * layout, parameter layout and the kernel launchers the host layer
 * (prr_net_host.cpp) uses. */
#ifndef PRR_NET_INTERNAL_H
#define PRR_NET_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prr_internal.h"

namespace prr {
namespace net {

enum {
    C = 32,                          /* channels of every activation */
    TAPS = 9,
    LAYER_W = TAPS * C * C,          /* W[tap][co][ci] */
    LAYER_PARAMS = LAYER_W + C,      /* then b[co] */
    LAYERS = 8,
    PARAMS = LAYERS * LAYER_PARAMS,
    OUT_CHANNELS = 12,               /* 2x2 sub-pixels x 3 colours */
};

/* Conv epilogues. */
enum {
    EPI_RELU = 0,        /* out = bf16(relu(v + b)) */
    EPI_RESID_RELU = 1,  /* r = bf16(relu(v + b)); out = bf16(resid + r); aux = r */
    EPI_LINEAR_F32 = 2,  /* out_f32 = v + b */
    EPI_LINEAR = 3,      /* out = bf16(v) */
    EPI_ADD = 4,         /* out = bf16(resid + v) */
};

/* One mosaic the network reads: code values, their conversion to the
 * normalized values develop works on, and the flips that bring its
 * pattern to RGGB. */
struct NetFrame {
    const uint16_t *samples;
    uint32_t stride, width, height;
    uint32_t curve[9];
    int32_t black;
    float one_over_range;
    float clip_r, clip_b;   /* red and blue sensels clipped at these */
    float feature;          /* 128 * sqrt(noise[0]) */
    /* The noise model of DemosaicParams (demosaic.rs:57-74): the variance
     * of a normalized value v is noise_c + max(noise_b, noise_a * (off + v)^2),
     * off = black / range; noise_a comes from the quantizer on the device
     * when qstatus is set. */
    float noise_a, noise_b, noise_c, noise_off;
    /* The green pass's offset in the variance: black * one_over_range
     * (green_row_impl, demosaic.rs:484), where the chroma pass uses
     * black / range (noise_off). */
    float gp_off;
    /* When set, the feature comes from the quantizer qmat0 * max_qscale
     * read on the device (develop(&frame, stage) after a device decode). */
    const prr_device_status *qstatus;
    uint32_t qmat0;
    uint32_t flip_y, flip_x;
    const float *target;    /* training: develop(Demosaiced) planes, width*height each */
};

/* A training crop: frame index and RGGB-order origin (even) of the
 * (2*H2) x (2*W2) window. */
struct Crop {
    uint32_t frame, y0, x0, pad_;
};

/* Compose parameters: white balance, clip and matrix after the network. */
struct Finish {
    float wb[2];
    int32_t has_matrix;
    float m[9];
};

/* Inference convolution modes (k_conv_infer) and the channels of the
 * network output the composition reads. */
enum { IM_HEAD = 0, IM_RES = 1, IM_TAIL = 2, OUT_CH = 12 };
hipError_t launch_conv_infer(int mode, const uint16_t *in, int h, int w, const uint16_t *wgt,
                             const float *bias, uint16_t *out, float *out12, const uint32_t *skip,
                             uint32_t grid, hipStream_t s);
const void *conv_infer_kernel();
/* Two consecutive layers (weights wa, ba then wb, bb) in one launch, the
 * first of them the network's first layer when `first` (else a residual
 * one), the second the last layer, writing out12, when `last` (else a
 * residual one, writing out): the same bits as two launch_conv_infer calls. */
hipError_t launch_conv_pair(int first, int last, const uint16_t *in, int h, int w,
                            const uint16_t *wa, const float *ba, const uint16_t *wb,
                            const float *bb, uint16_t *out, float *out12, const uint32_t *skip,
                            hipStream_t s);
hipError_t launch_weights(const float *params, uint16_t *wf, uint16_t *wt, hipStream_t s);
hipError_t launch_conv(int epi, const uint16_t *in, int n, int h, int w, const uint16_t *wgt,
                       const float *bias, const uint16_t *resid, uint16_t *out, uint16_t *aux,
                       float *out_f32, hipStream_t s);
size_t wgrad_tiles(int n, int h, int w);
hipError_t launch_wgrad(const uint16_t *x, const uint16_t *g, int n, int h, int w, float *part,
                        float *grad, hipStream_t s);
hipError_t launch_relu_grad(const uint16_t *dh, const uint16_t *mask, uint16_t *dz, size_t count,
                            hipStream_t s);
/* Adam at lr, decayed along a cosine to lr_min over `decay` steps (0 =
 * constant), the step count read on the device. */
hipError_t launch_adam(float *params, const float *grads, float *m, float *v, uint32_t *step,
                       float lr, float lr_min, uint32_t decay, hipStream_t s);
/* Training: crops[(step % steps) * n + k] (step read from *step) packed
 * into x0. */
hipError_t launch_pack_crops(const NetFrame *frames, const Crop *crops, const uint32_t *step,
                             uint32_t steps, int n, int h2, int w2, uint16_t *x0, hipStream_t s);
/* Mean over the loss cells of d^2 + d^4 * inv_tau2 (inv_tau2 = 0: the
 * squared error) and its gradient at out, rounded to bf16. */
hipError_t launch_loss(const NetFrame *frames, const Crop *crops, const uint32_t *step,
                       uint32_t steps, int n, int h2, int w2, int margin, float inv_tau2,
                       const float *out, uint16_t *dout, float *loss_part, float *loss,
                       hipStream_t s);
/* Inference over a whole frame (even width and height); nothing when
 * *skip is set. */
hipError_t launch_pack_frame(const NetFrame &f, uint16_t *x0, const uint32_t *skip, hipStream_t s);
/* The sampled-tile windows (WIN x WIN cells at most) packed as n images
 * of sy x sx cells for the training convolution. */
enum { WIN = 48 };
hipError_t launch_pack_windows(const NetFrame &f, const uint32_t *tiles, int n, int sy, int sx,
                               uint16_t *x0, hipStream_t s);
hipError_t launch_compose(const NetFrame &f, const Finish &fin, const float *out, float *r,
                          float *g, float *b, const uint32_t *skip, hipStream_t s);
/* Plain develop of sampled 64x64 tiles (tiles[2k], tiles[2k+1] = column
 * of 64, row of 64) into scratch (n_tiles * 3 * 4096 floats). */
hipError_t launch_develop_sampled(const DevelopDesc &d, const uint32_t *tiles, uint32_t n_tiles,
                                  float *scratch, hipStream_t s);
/* Sampled-tile check of the demosaic: the plain develop of each tile at
 * the stage's white balance and clip without the output matrix (d must
 * have has_matrix = 0) against the network's output on the tile's window
 * (wout: n_tiles images of sy x sx cells, 32 fp32 channels) finished the
 * same way; PSNR per tile into psnr, *use_plain set when any is below
 * min_db. */
hipError_t launch_check(const DevelopDesc &d, const NetFrame &f, const Finish &fin,
                        const float *wout, int sy, int sx, const uint32_t *tiles,
                        uint32_t n_tiles, float *scratch, float min_db, float *psnr,
                        uint32_t *use_plain, hipStream_t s);
hipError_t launch_gate_clear(uint32_t *flag, uint32_t value, hipStream_t s);
/* Sets *use_plain when the frame's noise feature lies outside [lo, hi]. */
hipError_t launch_noise_gate(const NetFrame &f, float lo, float hi, uint32_t *use_plain,
                             hipStream_t s);
hipError_t launch_step_inc(uint32_t *step, hipStream_t s);

} // namespace net
} // namespace prr

#endif
