/* Host side of the learned demosaic (include/proresraw_learned.h): device
 * buffers, the training step as a sequence of kernel launches (no
 * allocation, no host synchronization, so one step can be captured in a
 * graph), and inference with its on-device fallback to plain develop. */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/proresraw_learned.h"
#include "prr_net.h"

using namespace prr;
using namespace prr::net;

static_assert(PRR_NET_PARAMS == PARAMS, "parameter count");
static_assert(PRR_NET_LAYER_PARAMS == LAYER_PARAMS, "layer parameter count");

static int net_fail(prr_error *err, int code, const char *what, uint64_t a = 0, uint64_t b = 0)
{
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->code = code;
        if (what)
            std::strncpy(err->what, what, sizeof err->what - 1);
        err->a = a;
        err->b = b;
    }
    return code;
}

#define NET_TRY(call, what)                                                                      \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return net_fail(err, PRR_E_DEVICE, what, (uint64_t)e_);                              \
    } while (0)

namespace {
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    template <class T> T *as() const { return (T *)p; }
};
} // namespace

struct prr_net {
    prr_context *ctx = nullptr;
    int device = 0;
    float lo = 0.35f, hi = 1.06f;
    Buf params, m, v, grads, wf, wt, step;
    /* training */
    Buf frames, crops, x0, h[7], r[7], out, dout, dh[2], dz, part, loss_part, loss;
    uint32_t n_frames = 0, steps = 0, batch = 0, cells = 0, margin = 0;
    float lr_min = 0.0f;
    uint32_t decay = 0;
    float inv_tau2 = 0.0f; /* loss d^2 + d^4 / tau^2; 0: squared error */
    /* inference */
    Buf ix0, ia, ib, iout, scratch, psnr, flag;
    Buf wx0, wa, wb, wout; /* the sampled-tile windows */
    /* Sampled-tile lists per (width, height, count), uploaded on first use
     * so that later calls (and graph capture) copy nothing. */
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Buf> tiles;
    uint32_t last_k = 0;
    size_t last_cells = 0;  /* cells of the last whole-frame packing */
    bool pairs = true;      /* residual layers two a launch (k_conv_pair) */
    uint32_t conv_grid = 0; /* persistent inference convolution: workgroups */
};

static int grow(prr_net *net, Buf &b, size_t bytes, prr_error *err)
{
    if (b.p && b.bytes >= bytes)
        return PRR_OK;
    if (b.p) {
        NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
        NET_TRY(hipFree(b.p), "hipFree");
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    NET_TRY(hipMalloc(&b.p, want), "hipMalloc");
    b.bytes = want;
    return PRR_OK;
}

#define GROW(buf, bytes)                                                                         \
    do {                                                                                         \
        int rc_ = grow(net, buf, bytes, err);                                                    \
        if (rc_)                                                                                 \
            return rc_;                                                                          \
    } while (0)

extern "C" int prr_net_create(prr_context *ctx, prr_net **out, prr_error *err)
{
    if (!ctx || !out)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "ctx or out");
    *out = nullptr;
    prr_net *net = new prr_net();
    net->ctx = ctx;
    net->device = context_device(ctx);
    hipError_t e = hipSetDevice(net->device);
    int rc = e == hipSuccess ? PRR_OK : net_fail(err, PRR_E_DEVICE, "hipSetDevice", e);
    const size_t pb = sizeof(float) * PARAMS, wb = sizeof(uint16_t) * LAYERS * LAYER_W;
    if (!rc)
        rc = grow(net, net->params, pb, err);
    if (!rc)
        rc = grow(net, net->m, pb, err);
    if (!rc)
        rc = grow(net, net->v, pb, err);
    if (!rc)
        rc = grow(net, net->grads, pb, err);
    if (!rc)
        rc = grow(net, net->wf, wb, err);
    if (!rc)
        rc = grow(net, net->wt, wb, err);
    if (!rc)
        rc = grow(net, net->step, 64, err);
    if (!rc)
        rc = grow(net, net->flag, 64, err);
    if (!rc) {
        std::vector<float> zero(PARAMS, 0.0f);
        rc = prr_net_set_params(net, zero.data(), err);
    }
    if (rc) {
        prr_net_destroy(net);
        return rc;
    }
    *out = net;
    return PRR_OK;
}

extern "C" void prr_net_destroy(prr_net *net)
{
    if (!net)
        return;
    (void)hipSetDevice(net->device);
    (void)hipDeviceSynchronize();
    std::vector<Buf *> all = {&net->params, &net->m, &net->v, &net->grads, &net->wf, &net->wt,
                              &net->step, &net->frames, &net->crops, &net->x0, &net->out,
                              &net->dout, &net->dh[0], &net->dh[1], &net->dz, &net->part,
                              &net->loss_part, &net->loss, &net->ix0, &net->ia, &net->ib,
                              &net->iout, &net->scratch, &net->psnr, &net->flag,
                              &net->wx0, &net->wa, &net->wb, &net->wout};
    for (auto &kv : net->tiles)
        all.push_back(&kv.second);
    for (int l = 0; l < 7; l++) {
        all.push_back(&net->h[l]);
        all.push_back(&net->r[l]);
    }
    for (Buf *b : all)
        if (b->p)
            (void)hipFree(b->p);
    delete net;
}

extern "C" int prr_net_set_params(prr_net *net, const float *params, prr_error *err)
{
    if (!net || !params)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net or params");
    NET_TRY(hipSetDevice(net->device), "hipSetDevice");
    NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const size_t pb = sizeof(float) * PARAMS;
    NET_TRY(hipMemcpy(net->params.p, params, pb, hipMemcpyHostToDevice), "hipMemcpy params");
    std::vector<float> zero(PARAMS, 0.0f);
    NET_TRY(hipMemcpy(net->m.p, zero.data(), pb, hipMemcpyHostToDevice), "hipMemcpy m");
    NET_TRY(hipMemcpy(net->v.p, zero.data(), pb, hipMemcpyHostToDevice), "hipMemcpy v");
    NET_TRY(hipMemcpy(net->grads.p, zero.data(), pb, hipMemcpyHostToDevice), "hipMemcpy grads");
    const uint32_t step0 = 0;
    NET_TRY(hipMemcpy(net->step.p, &step0, 4, hipMemcpyHostToDevice), "hipMemcpy step");
    NET_TRY(launch_weights(net->params.as<float>(), net->wf.as<uint16_t>(), net->wt.as<uint16_t>(),
                           nullptr),
            "weights kernel");
    NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return PRR_OK;
}

extern "C" int prr_net_get_params(prr_net *net, float *params, prr_error *err)
{
    if (!net || !params)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net or params");
    NET_TRY(hipSetDevice(net->device), "hipSetDevice");
    NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    NET_TRY(hipMemcpy(params, net->params.p, sizeof(float) * PARAMS, hipMemcpyDeviceToHost),
            "hipMemcpy params");
    return PRR_OK;
}

extern "C" int prr_net_set_range(prr_net *net, float lo, float hi, prr_error *err)
{
    if (!net || !(lo <= hi))
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net or range");
    net->lo = lo;
    net->hi = hi;
    return PRR_OK;
}

/* The NetFrame of a mosaic developed at `d` (flips bring the pattern to
 * RGGB: the red sensel of BGGR sits at (1, 1), and so on). */
static NetFrame net_frame(const prr_header *hdr, const DevelopDesc &d,
                          const prr_demosaic_params &p, const prr_device_status *qstatus,
                          const float *target)
{
    NetFrame f;
    std::memset(&f, 0, sizeof f);
    f.samples = d.samples;
    f.stride = d.stride;
    f.width = d.width;
    f.height = d.height;
    for (int i = 0; i < 8; i++)
        f.curve[i] = d.curve[i];
    f.curve[8] = 65536u;
    f.black = d.black;
    f.one_over_range = d.one_over_range;
    f.clip_r = d.clip[0];
    f.clip_b = d.clip[1];
    f.feature = 128.0f * std::sqrt(p.noise[0]);
    f.noise_a = p.noise[0];
    f.noise_b = p.noise[1];
    f.noise_c = p.noise[2];
    f.noise_off = p.black_over_range;
    f.gp_off = (float)d.black * d.one_over_range;
    f.qstatus = qstatus;
    f.qmat0 = hdr->quant_matrix[0];
    f.flip_y = (uint32_t)(d.cfa_phase >> 1) & 1u;
    f.flip_x = (uint32_t)d.cfa_phase & 1u;
    f.target = target;
    return f;
}

extern "C" int prr_net_train_setup(prr_net *net, const prr_net_frame *frames, uint32_t n_frames,
                                   const uint32_t *crops, uint32_t steps, uint32_t batch,
                                   uint32_t cells, uint32_t margin, prr_error *err)
{
    if (!net || !frames || !crops || !n_frames || !steps || !batch || !cells ||
        2 * margin >= cells)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "training setup");
    NET_TRY(hipSetDevice(net->device), "hipSetDevice");
    std::vector<NetFrame> nf(n_frames);
    for (uint32_t i = 0; i < n_frames; i++) {
        const prr_net_frame &fr = frames[i];
        if (!fr.d_samples || !fr.d_target)
            return net_fail(err, PRR_E_INVALID_ARGUMENT, "frame samples or target", i);
        DevelopDesc d = develop_desc(&fr.header, fr.d_samples, &fr.params, PRR_STAGE_DEMOSAICED);
        d.stride = fr.stride;
        nf[i] = net_frame(&fr.header, d, fr.params, nullptr, fr.d_target);
        if (nf[i].flip_x || nf[i].flip_y)
            return net_fail(err, PRR_E_INVALID_ARGUMENT, "training frames must be RGGB", i);
    }
    std::vector<Crop> cr((size_t)steps * batch);
    for (size_t k = 0; k < cr.size(); k++) {
        cr[k].frame = crops[3 * k];
        cr[k].y0 = crops[3 * k + 1];
        cr[k].x0 = crops[3 * k + 2];
        cr[k].pad_ = 0;
        if (cr[k].frame >= n_frames || (cr[k].y0 | cr[k].x0) & 1u ||
            cr[k].y0 + 2 * cells > nf[cr[k].frame].height ||
            cr[k].x0 + 2 * cells > nf[cr[k].frame].width)
            return net_fail(err, PRR_E_INVALID_ARGUMENT, "crop outside its frame", k);
    }
    const size_t act = (size_t)batch * cells * cells * C;
    GROW(net->frames, sizeof(NetFrame) * n_frames);
    GROW(net->crops, sizeof(Crop) * cr.size());
    GROW(net->x0, 2 * act);
    for (int l = 0; l < 7; l++) {
        GROW(net->h[l], 2 * act);
        if (l)
            GROW(net->r[l], 2 * act);
    }
    GROW(net->out, 4 * act);
    GROW(net->dout, 2 * act);
    GROW(net->dh[0], 2 * act);
    GROW(net->dh[1], 2 * act);
    GROW(net->dz, 2 * act);
    GROW(net->part, sizeof(float) * LAYER_PARAMS * wgrad_tiles((int)batch, (int)cells, (int)cells));
    GROW(net->loss_part, sizeof(float) * ((act / C + 255) / 256 + 1));
    GROW(net->loss, 64);
    NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    NET_TRY(hipMemcpy(net->frames.p, nf.data(), sizeof(NetFrame) * n_frames, hipMemcpyHostToDevice),
            "hipMemcpy frames");
    NET_TRY(hipMemcpy(net->crops.p, cr.data(), sizeof(Crop) * cr.size(), hipMemcpyHostToDevice),
            "hipMemcpy crops");
    net->n_frames = n_frames;
    net->steps = steps;
    net->batch = batch;
    net->cells = cells;
    net->margin = margin;
    return PRR_OK;
}

extern "C" int prr_net_train_step(prr_net *net, float lr, void *stream, prr_error *err)
{
    if (!net || !net->steps)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "no training setup");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)net->batch, hw = (int)net->cells;
    const size_t count = (size_t)n * hw * hw * C;
    uint16_t *wf = net->wf.as<uint16_t>(), *wt = net->wt.as<uint16_t>();
    float *params = net->params.as<float>(), *grads = net->grads.as<float>();
    uint32_t *step = net->step.as<uint32_t>();
    auto W = [&](int l) { return wf + (size_t)l * LAYER_W; };
    auto WT = [&](int l) { return wt + (size_t)l * LAYER_W; };
    auto B = [&](int l) { return params + (size_t)l * LAYER_PARAMS + LAYER_W; };
    auto G = [&](int l) { return grads + (size_t)l * LAYER_PARAMS; };
    auto H = [&](int l) { return net->h[l].as<uint16_t>(); };
    float *part = net->part.as<float>();

    NET_TRY(launch_pack_crops(net->frames.as<NetFrame>(), net->crops.as<Crop>(), step,
                              net->steps, n, hw, hw, net->x0.as<uint16_t>(), s),
            "pack kernel");
    NET_TRY(launch_conv(EPI_RELU, net->x0.as<uint16_t>(), n, hw, hw, W(0), B(0), nullptr, H(0),
                        nullptr, nullptr, s),
            "conv kernel");
    for (int l = 1; l < 7; l++)
        NET_TRY(launch_conv(EPI_RESID_RELU, H(l - 1), n, hw, hw, W(l), B(l), H(l - 1), H(l),
                            net->r[l].as<uint16_t>(), nullptr, s),
                "conv kernel");
    NET_TRY(launch_conv(EPI_LINEAR_F32, H(6), n, hw, hw, W(7), B(7), nullptr, nullptr, nullptr,
                        net->out.as<float>(), s),
            "conv kernel");
    NET_TRY(launch_loss(net->frames.as<NetFrame>(), net->crops.as<Crop>(), step, net->steps, n,
                        hw, hw, (int)net->margin, net->inv_tau2, net->out.as<float>(),
                        net->dout.as<uint16_t>(), net->loss_part.as<float>(),
                        net->loss.as<float>(), s),
            "loss kernel");

    NET_TRY(launch_wgrad(H(6), net->dout.as<uint16_t>(), n, hw, hw, part, G(7), s), "wgrad");
    int a = 0;
    NET_TRY(launch_conv(EPI_LINEAR, net->dout.as<uint16_t>(), n, hw, hw, WT(7), nullptr, nullptr,
                        net->dh[a].as<uint16_t>(), nullptr, nullptr, s),
            "dgrad kernel");
    uint16_t *dz = net->dz.as<uint16_t>();
    for (int l = 6; l >= 1; l--) {
        NET_TRY(launch_relu_grad(net->dh[a].as<uint16_t>(), net->r[l].as<uint16_t>(), dz, count,
                                 s),
                "relu grad");
        NET_TRY(launch_wgrad(H(l - 1), dz, n, hw, hw, part, G(l), s), "wgrad");
        NET_TRY(launch_conv(EPI_ADD, dz, n, hw, hw, WT(l), nullptr, net->dh[a].as<uint16_t>(),
                            net->dh[1 - a].as<uint16_t>(), nullptr, nullptr, s),
                "dgrad kernel");
        a = 1 - a;
    }
    NET_TRY(launch_relu_grad(net->dh[a].as<uint16_t>(), H(0), dz, count, s), "relu grad");
    NET_TRY(launch_wgrad(net->x0.as<uint16_t>(), dz, n, hw, hw, part, G(0), s), "wgrad");
    if (lr > 0.0f) {
        NET_TRY(launch_adam(params, grads, net->m.as<float>(), net->v.as<float>(), step, lr,
                            net->lr_min, net->decay, s),
                "adam kernel");
        NET_TRY(launch_weights(params, wf, wt, s), "weights kernel");
    }
    NET_TRY(launch_step_inc(step, s), "step kernel");
    return PRR_OK;
}

extern "C" int prr_net_set_schedule(prr_net *net, float lr_min, uint32_t decay_steps,
                                    prr_error *err)
{
    if (!net || !(lr_min >= 0.0f))
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net or lr_min");
    net->lr_min = lr_min;
    net->decay = decay_steps;
    return PRR_OK;
}

extern "C" int prr_net_set_loss(prr_net *net, float tau, prr_error *err)
{
    if (!net || !(tau >= 0.0f) || tau > 1e18f)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net or tau");
    net->inv_tau2 = tau > 0.0f ? 1.0f / (tau * tau) : 0.0f;
    return PRR_OK;
}

extern "C" int prr_net_read(prr_net *net, int what, uint32_t layer, void *host, size_t bytes,
                            prr_error *err)
{
    if (!net || !host)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net or host");
    const size_t act = (size_t)net->batch * net->cells * net->cells * C;
    const void *src = nullptr;
    size_t need = 0;
    switch (what) {
    case PRR_NET_X0: src = net->x0.p; need = 2 * act; break;
    case PRR_NET_H:
        if (layer > 6)
            return net_fail(err, PRR_E_INVALID_ARGUMENT, "layer", layer);
        src = net->h[layer].p;
        need = 2 * act;
        break;
    case PRR_NET_R:
        if (layer < 1 || layer > 6)
            return net_fail(err, PRR_E_INVALID_ARGUMENT, "layer", layer);
        src = net->r[layer].p;
        need = 2 * act;
        break;
    case PRR_NET_OUT: src = net->out.p; need = 4 * act; break;
    case PRR_NET_DOUT: src = net->dout.p; need = 2 * act; break;
    case PRR_NET_GRAD: src = net->grads.p; need = sizeof(float) * PARAMS; break;
    case PRR_NET_LOSS: src = net->loss.p; need = 4; break;
    case PRR_NET_STEP: src = net->step.p; need = 4; break;
    case PRR_NET_FRAME_X0: src = net->last_cells ? net->ix0.p : nullptr;
        need = 2 * net->last_cells * C;
        break;
    default: return net_fail(err, PRR_E_INVALID_ARGUMENT, "what", (uint64_t)what);
    }
    if (!src)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "no such tensor yet");
    if (bytes != need)
        return net_fail(err, PRR_E_OUTPUT_SIZE, "tensor bytes", need, bytes);
    NET_TRY(hipSetDevice(net->device), "hipSetDevice");
    NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    NET_TRY(hipMemcpy(host, src, need, hipMemcpyDeviceToHost), "hipMemcpy");
    return PRR_OK;
}

/* Sampled-tile positions: K tiles of 64x64 spread over the frame on a
 * jittered diagonal, (column, row) in units of 64. */
static std::vector<uint32_t> sample_tiles(uint32_t w, uint32_t h, uint32_t k)
{
    const uint32_t nbx = (w + 63) / 64, nby = (h + 63) / 64;
    std::vector<uint32_t> t(2 * (size_t)k);
    for (uint32_t i = 0; i < k; i++) {
        t[2 * i] = std::min(nbx - 1, (uint32_t)(((2ull * i + 1) * nbx) / (2ull * k)));
        const uint32_t j = (i * 5 + 2) % k;
        t[2 * i + 1] = std::min(nby - 1, (uint32_t)(((2ull * j + 1) * nby) / (2ull * k)));
    }
    return t;
}

extern "C" int prr_net_develop_device(prr_net *net, const prr_header *hdr,
                                      const uint16_t *d_samples,
                                      const prr_demosaic_params *params,
                                      const prr_device_status *d_status, int stage,
                                      uint32_t check_tiles, float min_db, float *d_r, float *d_g,
                                      float *d_b, void *stream, prr_error *err)
{
    if (!net || !hdr || !d_samples || !params || !d_r || !d_g || !d_b)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "null argument");
    if (stage < PRR_STAGE_DEMOSAICED || stage > PRR_STAGE_REC2020)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "stage", (uint64_t)stage);
    const uint32_t w = hdr->width, h = hdr->height;
    if (!w || !h || (w & 1) || (h & 1))
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "width and height must be even", w, h);
    hipStream_t s = (hipStream_t)stream;
    NET_TRY(hipSetDevice(net->device), "hipSetDevice");
    DevelopDesc d = develop_desc(hdr, d_samples, params, stage);
    if (d_status) {
        d.qstatus = d_status;
        d.qmat0 = hdr->quant_matrix[0];
        d.range = (uint16_t)(hdr->white_level - 256u);
    }
    const NetFrame f = net_frame(hdr, d, *params, d_status, nullptr);
    Finish fin;
    std::memset(&fin, 0, sizeof fin);
    fin.wb[0] = d.wb[0];
    fin.wb[1] = d.wb[1];
    fin.has_matrix = d.has_matrix;
    std::memcpy(fin.m, d.m, sizeof fin.m);

    const size_t cells = (size_t)(w / 2) * (h / 2);
    GROW(net->ix0, 2 * cells * C);
    GROW(net->ia, 2 * cells * C);
    GROW(net->ib, 2 * cells * C);
    GROW(net->iout, 4 * cells * OUT_CH);
    if (!net->conv_grid) {
        hipDeviceProp_t prop;
        NET_TRY(hipGetDeviceProperties(&prop, net->device), "hipGetDeviceProperties");
        int per_cu = 0;
        NET_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, conv_infer_kernel(), 256, 0),
                "hipOccupancy");
        net->conv_grid = (uint32_t)(prop.multiProcessorCount * std::max(per_cu, 1));
    }
    const uint32_t k = std::min<uint32_t>(check_tiles, 64);
    GROW(net->scratch, sizeof(float) * 3 * 4096 * 64);
    GROW(net->psnr, sizeof(float) * 64);
    const uint32_t *d_tiles = nullptr;
    if (k) {
        Buf &tb = net->tiles[std::make_tuple(w, h, k)];
        if (!tb.p) {
            std::vector<uint32_t> t = sample_tiles(w, h, k);
            GROW(tb, sizeof(uint32_t) * t.size());
            NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
            NET_TRY(hipMemcpy(tb.p, t.data(), sizeof(uint32_t) * t.size(), hipMemcpyHostToDevice),
                    "hipMemcpy tiles");
        }
        d_tiles = tb.as<uint32_t>();
    }
    uint32_t *flag = net->flag.as<uint32_t>();
    const uint16_t *wf = net->wf.as<uint16_t>();
    const float *params_d = net->params.as<float>();
    auto W = [&](int l) { return wf + (size_t)l * LAYER_W; };
    auto B = [&](int l) { return params_d + (size_t)l * LAYER_PARAMS + LAYER_W; };
    const int hh = (int)h / 2, ww = (int)w / 2;

    NET_TRY(launch_gate_clear(flag, 0u, s), "gate kernel");
    NET_TRY(launch_noise_gate(f, net->lo, net->hi, flag, s), "gate kernel");
    /* The check first, on the network run over the sampled tiles' windows
     * (the training convolution, the same arithmetic as the whole-frame
     * one), so that a frame that falls back never runs the whole network. */
    if (k) {
        const int sy = std::min(hh, (int)WIN), sx = std::min(ww, (int)WIN);
        const size_t wcells = (size_t)k * sy * sx;
        GROW(net->wx0, 2 * wcells * C);
        GROW(net->wa, 2 * wcells * C);
        GROW(net->wb, 2 * wcells * C);
        GROW(net->wout, 4 * wcells * C);
        NET_TRY(launch_pack_windows(f, d_tiles, (int)k, sy, sx, net->wx0.as<uint16_t>(), s),
                "pack kernel");
        uint16_t *a = net->wa.as<uint16_t>(), *b = net->wb.as<uint16_t>();
        NET_TRY(launch_conv(EPI_RELU, net->wx0.as<uint16_t>(), (int)k, sy, sx, W(0), B(0), nullptr,
                            a, nullptr, nullptr, s),
                "conv kernel");
        for (int l = 1; l < 7; l++) {
            NET_TRY(launch_conv(EPI_RESID_RELU, a, (int)k, sy, sx, W(l), B(l), a, b, nullptr,
                                nullptr, s),
                    "conv kernel");
            std::swap(a, b);
        }
        NET_TRY(launch_conv(EPI_LINEAR_F32, a, (int)k, sy, sx, W(7), B(7), nullptr, nullptr,
                            nullptr, net->wout.as<float>(), s),
                "conv kernel");
        DevelopDesc dc = d;
        dc.has_matrix = 0;
        NET_TRY(launch_check(dc, f, fin, net->wout.as<float>(), sy, sx, d_tiles, k,
                             net->scratch.as<float>(), min_db, net->psnr.as<float>(), flag, s),
                "check kernels");
    }
    /* The whole frame through the network when it did not fall back. */
    NET_TRY(launch_pack_frame(f, net->ix0.as<uint16_t>(), flag, s), "pack kernel");
    uint16_t *a = net->ia.as<uint16_t>(), *b = net->ib.as<uint16_t>();
    const uint32_t g = net->conv_grid;
    if (net->pairs) {
        /* Layers (0, 1), (2, 3), (4, 5), (6, 7), two a launch. */
        const uint16_t *in = net->ix0.as<uint16_t>();
        for (int l = 0; l < 8; l += 2) {
            NET_TRY(launch_conv_pair(l == 0, l == 6, in, hh, ww, W(l), B(l), W(l + 1), B(l + 1), a,
                                     net->iout.as<float>(), flag, s),
                    "conv kernel");
            in = a;
            std::swap(a, b);
        }
    } else {
        NET_TRY(launch_conv_infer(IM_HEAD, net->ix0.as<uint16_t>(), hh, ww, W(0), B(0), a,
                                  nullptr, flag, g, s),
                "conv kernel");
        for (int l = 1; l < 7; l++) {
            NET_TRY(launch_conv_infer(IM_RES, a, hh, ww, W(l), B(l), b, nullptr, flag, g, s),
                    "conv kernel");
            std::swap(a, b);
        }
        NET_TRY(launch_conv_infer(IM_TAIL, a, hh, ww, W(7), B(7), nullptr, net->iout.as<float>(),
                                  flag, g, s),
                "conv kernel");
    }
    NET_TRY(launch_compose(f, fin, net->iout.as<float>(), d_r, d_g, d_b, flag, s),
            "compose kernel");
    d.gate = flag;
    NET_TRY(launch_develop(d, d_r, d_g, d_b, s), "develop kernel");
    net->last_k = k;
    net->last_cells = cells;
    return PRR_OK;
}

extern "C" int prr_net_set_pairs(prr_net *net, int on, prr_error *err)
{
    if (!net)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net");
    net->pairs = on != 0;
    return PRR_OK;
}

extern "C" int prr_net_last_check(prr_net *net, uint32_t *used_plain, float *min_psnr,
                                  prr_error *err)
{
    if (!net)
        return net_fail(err, PRR_E_INVALID_ARGUMENT, "net");
    NET_TRY(hipSetDevice(net->device), "hipSetDevice");
    NET_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    uint32_t flag = 0;
    NET_TRY(hipMemcpy(&flag, net->flag.p, 4, hipMemcpyDeviceToHost), "hipMemcpy flag");
    if (used_plain)
        *used_plain = flag;
    if (min_psnr) {
        float lo = INFINITY;
        if (net->last_k && net->psnr.p) {
            std::vector<float> p(net->last_k);
            NET_TRY(hipMemcpy(p.data(), net->psnr.p, sizeof(float) * p.size(),
                              hipMemcpyDeviceToHost),
                    "hipMemcpy psnr");
            for (float v : p)
                lo = std::min(lo, v);
        }
        *min_psnr = lo;
    }
    return PRR_OK;
}
