/* proresraw_amd learned demosaic: a small convolutional demosaic network
 * (bf16 weights on the matrix cores, fp32 accumulation) that stands beside
 * the plain develop of proresraw_amd.h.
 *
 * No function here has a counterpart in the proresraw_rs crate: the crate
 * has only the plain demosaic (src/demosaic.rs). The network is trained on
 * pairs of mosaics and develop(Stage::Demosaiced) output (demosaic.rs:190);
 * the plain kernel stays the default drop-in, the judge of the network's
 * error, and the fallback whenever a frame's noise level is outside the
 * trained range or a sampled-tile check falls below the PSNR floor.
 *
 * Conventions are those of proresraw_amd.h: 0 or a PRR_E_* status, errors
 * in `err`, device pointers and a hipStream_t as is for *_device calls.
 */
#ifndef PRORESRAW_LEARNED_H
#define PRORESRAW_LEARNED_H

#include "proresraw_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 8 layers of W[tap 9][out 32][in 32] followed by bias[32]: fp32. */
#define PRR_NET_PARAMS 73984
#define PRR_NET_LAYER_PARAMS 9248

typedef struct prr_net prr_net;

/* A mosaic the network reads: device code values with their header and
 * develop parameters (black, range, noise, clip), and for training the
 * develop(Stage::Demosaiced) planes, width*height floats each. */
typedef struct {
    const uint16_t *d_samples;
    uint32_t stride;
    prr_header header;
    prr_demosaic_params params;
    const float *d_target;
} prr_net_frame;

int prr_net_create(prr_context *ctx, prr_net **out, prr_error *err);
void prr_net_destroy(prr_net *net);

/* Parameters, PRR_NET_PARAMS host floats. Setting them resets the Adam
 * state and step count. */
int prr_net_set_params(prr_net *net, const float *params, prr_error *err);
int prr_net_get_params(prr_net *net, float *params, prr_error *err);
/* The noise-feature range the parameters were trained on (feature =
 * 128 * sqrt(noise[0])); frames outside it take the plain develop. */
int prr_net_set_range(prr_net *net, float feature_lo, float feature_hi, prr_error *err);

/* ---- training ---------------------------------------------------------- */
/* Frames and crops for the steps to come: crops[3 * (step * batch + k)] =
 * frame index, top, left (even) of a window of 2*cells x 2*cells sensels;
 * the loss covers its cells at least `margin` from the edge. Uploaded once;
 * each prr_net_train_step reads its crops with the step count kept on the
 * device (modulo `steps`), so a captured graph of one step replays the
 * run. */
int prr_net_train_setup(prr_net *net, const prr_net_frame *frames, uint32_t n_frames,
                        const uint32_t *crops, uint32_t steps, uint32_t batch, uint32_t cells,
                        uint32_t margin, prr_error *err);
/* Learning-rate schedule of the following steps: the rate given to
 * prr_net_train_step decays along a cosine to lr_min over decay_steps
 * steps (0, the default, keeps it constant); computed on the device from
 * the step count. */
int prr_net_set_schedule(prr_net *net, float lr_min, uint32_t decay_steps, prr_error *err);
/* Loss of the following steps: the mean of d^2 + d^4 / tau^2 over the loss
 * cells' colours (d = network + bilinear - target), which weighs the few
 * large errors on edges more than the squared error does; tau = 0 (the
 * default) is the squared error. Takes effect in steps enqueued (or
 * captured) afterwards. */
int prr_net_set_loss(prr_net *net, float tau, prr_error *err);
/* One step: pack, forward, loss, backward, Adam (lr) unless lr is 0, step
 * count + 1. Enqueued on `stream`; allocates nothing. */
int prr_net_train_step(prr_net *net, float lr, void *stream, prr_error *err);

/* Tensors of the last step, for checking (synchronous copies to host):
 * PRR_NET_X0 / H (layer 0..6) / R (layer 1..6) / DOUT: bf16, batch*cells*
 * cells*32 values; OUT: fp32, same count; GRAD: fp32, PRR_NET_PARAMS;
 * LOSS: one fp32; STEP: one uint32. PRR_NET_FRAME_X0: the whole-frame input
 * packed by the last prr_net_develop_device that ran the network, bf16,
 * (height/2)*(width/2)*32 values. */
enum { PRR_NET_X0 = 0, PRR_NET_H = 1, PRR_NET_R = 2, PRR_NET_OUT = 3, PRR_NET_DOUT = 4,
       PRR_NET_GRAD = 5, PRR_NET_LOSS = 6, PRR_NET_STEP = 7, PRR_NET_FRAME_X0 = 8 };
int prr_net_read(prr_net *net, int what, uint32_t layer, void *host, size_t bytes,
                 prr_error *err);

/* ---- inference ----------------------------------------------------------- */
/* develop_planar_into's counterpart through the network: the mosaic of
 * d_samples (even width and height) developed at `stage` into three planes
 * of width*height floats. The noise feature comes from `params`, or from
 * d_status (its max_qscale) when that is not NULL. The output is the plain
 * develop when the feature is outside the trained range, or when any of
 * `check_tiles` sampled 64x64 tiles (0 = none) has a PSNR (peak 1) below
 * min_db between the network's demosaic and the plain one, both white
 * balanced and clipped as the stage says but before the output matrix
 * (which is exact and the same for both); the choice is made on the
 * device. */
int prr_net_develop_device(prr_net *net, const prr_header *hdr, const uint16_t *d_samples,
                           const prr_demosaic_params *params, const prr_device_status *d_status,
                           int stage, uint32_t check_tiles, float min_db, float *d_r, float *d_g,
                           float *d_b, void *stream, prr_error *err);
/* Whether prr_net_develop_device runs the network's eight layers two a
 * launch (on: the default; each pair keeps its middle activations in LDS and
 * reads and writes HBM once) or one a launch. Both give the same bits. */
int prr_net_set_pairs(prr_net *net, int on, prr_error *err);

/* After the stream has run the last prr_net_develop_device: whether it
 * fell back to plain develop, and the lowest sampled-tile PSNR (dB).
 * A prr_net holds one set of activation, check and flag buffers: calls on
 * it must be ordered (one stream), or they race. */
int prr_net_last_check(prr_net *net, uint32_t *used_plain, float *min_psnr, prr_error *err);

#ifdef __cplusplus
}
#endif

#endif
