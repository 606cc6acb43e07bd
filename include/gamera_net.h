/* gamera_net.h -- C-ABI of libgamera_net.so: the learned coarse-grid
 * correction network (no counterpart in the reference).
 *
 * A 3-layer 3x3x3 convolution net on the ghost-inclusive cell grid of a block
 * (8 input features -> 32 -> 32 -> 8 outputs, leaky ReLU x > 0 ? x : x/8 between), BF16 weights and
 * activations, FP32 accumulation on MFMA.  Inputs: rho, v, p, B of a state,
 * divided by fixed scales.  Outputs per cell: 5 gas-flux and 3 edge-EMF
 * corrections, handed to the step with gam_set_correction (gamera_hip.h).
 *
 * Parameters are FP32 masters, layer by layer: W[27][Co][Ci] then b[Co], with
 * (Co, Ci) = (32, 8), (32, 32), (8, 32); tap = dz*9 + dy*3 + dx, d = -1..1 -> 0..2.
 * All device pointers are the caller's; the net owns its activations. */
#ifndef GAMERA_NET_H
#define GAMERA_NET_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gam_net gam_net;

int  gam_net_num_params(void);
/* Ni, Nj, Nk: the ghost-inclusive cell extents of the block (each >= 8) */
int  gam_net_create(int Ni, int Nj, int Nk, gam_net **out);
int  gam_net_destroy(gam_net *n);
const char *gam_net_error(const gam_net *n);
void *gam_net_stream(gam_net *n);
/* Queue the net's work on the caller's hipStream_t (e.g. gam_stream(h) of the
 * block it corrects, so inference, the step and training need no host
 * synchronisation between them); NULL goes back to the net's own stream.  The
 * caller's stream must outlive that use. */
int  gam_net_set_stream(gam_net *n, void *stream);
int  gam_net_sync(gam_net *n);

int  gam_net_set_params(gam_net *n, const float *w);               /* host FP32 [num_params] */
int  gam_net_get_params(gam_net *n, float *w, float *grad);        /* host; NULL skips */
int  gam_net_set_scales(gam_net *n, float rho, float v, float p, float b, float gamma);
/* features: (rho - rho0)/rho, v/v, (p - p0)/p, B/b with the scales above (offsets default 0) */
int  gam_net_set_offsets(gam_net *n, float rho0, float p0);

/* Features from a device state (Gas G*5, Bxyz G*3; precision 0 FP64, 1 FP32),
 * or from host FP32 features [G][8] (tests). */
int  gam_net_features(gam_net *n, const void *Gas, const void *Bxyz, int precision);
int  gam_net_set_features(gam_net *n, const float *feat);

int  gam_net_forward(gam_net *n);
/* Inference only: features of the device state and the three layers in one
 * call, each layer's input staged in LDS per brick.  The output is stored on
 * the cells at least `halo` (>= 3) cells from the grid faces and equals
 * gam_net_forward's there bit for bit; elsewhere it keeps its old values, and
 * the activations kept for gam_net_backward are not all rewritten (train
 * through gam_net_features + gam_net_forward). */
int  gam_net_infer(gam_net *n, const void *Gas, const void *Bxyz, int precision, int halo);
/* The hipEvent_t gam_net_infer records after its last kernel (the net's
 * lifetime): hand it to gam_set_correction_wait so that a step on another
 * stream waits for the output only where it reads it. */
void *gam_net_infer_event(gam_net *n);
const float *gam_net_output(gam_net *n);                           /* device FP32 [G][8] */
int  gam_net_get_output(gam_net *n, float *out);                   /* host copy */

/* Accumulate dL/dparams for the upstream gradient dout (device FP32 [G][8])
 * through the activations of the last forward. */
int  gam_net_zero_grad(gam_net *n);
int  gam_net_backward(gam_net *n, const float *dout);
int  gam_net_backward_host(gam_net *n, const float *dout);         /* dout from host memory */
/* a device FP32 [G][8] buffer owned by the net, for the caller to write an
 * upstream gradient into (e.g. with gam_correction_grad) before gam_net_backward */
float *gam_net_grad_buffer(gam_net *n);
/* Adam step t (>= 1) on the accumulated gradient */
int  gam_net_adam(gam_net *n, float lr, float b1, float b2, float eps, int t);
/* BF16 activations [G][32] of the last forward (layer 0 features, 1, 2 hidden); tests */
int  gam_net_get_act(gam_net *n, int layer, unsigned short *out);

#ifdef __cplusplus
}
#endif
#endif
