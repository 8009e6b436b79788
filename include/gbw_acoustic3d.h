/*
 * gbw_acoustic3d.h -- C ABI of the MI355X (gfx950) 3D acoustic propagator.
 *
 * Replaces the time stepping of GeoBrain's 3D acoustic forward kernel:
 *   acoustic_forward_kernel_3d(nx, ny, nz, dx, dy, dz, nt, dt, nabc,
 *                              free_surface, src_x, src_y, src_z, src_n, src_v,
 *                              rcv_x, rcv_y, rcv_z, rcv_n, damp, vp, rho,
 *                              checkpoint_segments, device, dtype)
 *       geobrain/physics/wave/kernels/acoustic_3d.py:275-476
 * and the per-step body it drives,
 *   acoustic_step_forward_3d(...)   geobrain/physics/wave/kernels/acoustic_3d.py:20-273
 * Forward modelling (seismograms, energy maps, checkpoints) and the gradient
 * w.r.t. vp, rho and src_v (the hand-written adjoint of the step, recomputing
 * each checkpoint interval).  Conventions are those of gbw_acoustic2d.h: 0 or
 * a negative gbw_status, gbw_last_error() for the message, DEVICE / HOST
 * pointers, no allocation (the caller passes a workspace), work on the
 * caller's stream.
 */
#ifndef GBW_ACOUSTIC3D_H
#define GBW_ACOUSTIC3D_H

#include <stddef.h>
#include <stdint.h>

#include "gbw_acoustic2d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scalars of acoustic_forward_kernel_3d (acoustic_3d.py:275-301) plus a launch
 * choice.  dx and dy are accepted for signature parity; the scheme uses dz on
 * every axis, as the reference does (acoustic_3d.py:403,407). */
typedef struct gbw_acoustic3d_problem {
  int32_t nx, ny, nz;        /* model grid (unpadded)                                  */
  float dx, dy, dz;          /* grid spacing (m)                                       */
  int32_t nt;                /* time steps                                             */
  float dt;                  /* time step (s)                                          */
  int32_t nabc;              /* absorbing layers per side (>= 1)                       */
  int32_t free_surface;      /* 0 / 1                                                  */
  int32_t src_n;             /* shots                                                  */
  int32_t rcv_n;             /* receivers shared by every shot                         */
  int32_t wavelet_shared;    /* 1: src_v is (nt,) and feeds every shot; 0: (src_n, nt) */
                             /* (every point of a shot); 2: (src_n, src_points, nt)     */
  int32_t energy_chunk;      /* steps per chunk whose energy sums are added (0 = nt)   */
  int32_t shots_per_block;   /* launch choice: shots per workgroup (0 = all); 1, 2 or 4
                                with the z-marching kernel (0 = 4); 1 to 8 with the
                                fused step (0 = min(src_n, 8))                         */
  int32_t checkpoint_every;  /* steps per checkpoint interval (0 = no checkpoints)     */
  int32_t step_kernel;       /* launch choice: 0 separate p and v step kernels (default),
                                1 fused z-marching step (tiles in workgroups),
                                2 fused step with rows in wavefronts (forward pass;
                                the gradient's recompute keeps the p and v kernels)   */
  int32_t z_chunk;           /* launch choice: z planes per z-marching workgroup (0 = auto) */
  int32_t src_points;        /* point sources per shot (0 or 1: one; K > 1: src_x, src_y
                                and src_z are (src_n, K), distinct cells within a shot,
                                as the reference's 2-D positions, acoustic_3d.py:171-176;
                                such problems run the separate p and v step kernels) */
} gbw_acoustic3d_problem;

/* Geometry and model (acoustic_3d.py:286-297); positions are unpadded indices. */
typedef struct gbw_acoustic3d_inputs {
  const int64_t* src_x;      /* HOST (src_n,), or (src_n, src_points) row-major        */
  const int64_t* src_y;      /* HOST (src_n,), or (src_n, src_points) row-major        */
  const int64_t* src_z;      /* HOST (src_n,), or (src_n, src_points) row-major        */
  const int64_t* rcv_x;      /* HOST (rcv_n,)                                          */
  const int64_t* rcv_y;      /* HOST (rcv_n,)                                          */
  const int64_t* rcv_z;      /* HOST (rcv_n,)                                          */
  const float* src_v;        /* DEVICE (src_n, nt), (nt,) or (src_n, src_points, nt)   */
  const float* damp;         /* DEVICE (nz + 2 nabc, ny + 2 nabc, nx + 2 nabc)         */
  const float* vp;           /* DEVICE (nz, ny, nx)                                    */
  const float* rho;          /* DEVICE (nz, ny, nx)                                    */
} gbw_acoustic3d_inputs;

/* Outputs (the dict of acoustic_3d.py:466-476); any pointer may be NULL. */
typedef struct gbw_acoustic3d_outputs {
  float* rec_p;              /* DEVICE (src_n, nt, rcv_n)  key 'p'                     */
  float* rec_vx;             /* DEVICE (src_n, nt, rcv_n)  key 'vx'                    */
  float* rec_vy;             /* DEVICE (src_n, nt, rcv_n)  key 'vy'                    */
  float* rec_vz;             /* DEVICE (src_n, nt, rcv_n)  key 'vz'                    */
  float* energy_p;           /* DEVICE (nz, ny, nx)  key 'forward_wavefield_p'         */
  float* energy_vx;          /* DEVICE (nz, ny, nx)  key 'forward_wavefield_vx'        */
  float* energy_vy;          /* DEVICE (nz, ny, nx)  key 'forward_wavefield_vy'        */
  float* energy_vz;          /* DEVICE (nz, ny, nx)  key 'forward_wavefield_vz'        */
  void* checkpoints;         /* DEVICE checkpoint_bytes(), needed when checkpoint_every */
} gbw_acoustic3d_outputs;

/* Gradient request: adjoint sources are dLoss/d(seismogram).  Any g_* may be
 * NULL (zero); any grad_* may be NULL (not wanted). */
typedef struct gbw_acoustic3d_adjoint {
  const float* g_p;          /* DEVICE (src_n, nt, rcv_n)                              */
  const float* g_vx;         /* DEVICE (src_n, nt, rcv_n)                              */
  const float* g_vy;         /* DEVICE (src_n, nt, rcv_n)                              */
  const float* g_vz;         /* DEVICE (src_n, nt, rcv_n)                              */
  const void* checkpoints;   /* DEVICE buffer filled by gbw_acoustic3d_forward         */
  float* grad_vp;            /* DEVICE (nz, ny, nx)   overwritten                      */
  float* grad_rho;           /* DEVICE (nz, ny, nx)   overwritten; density as an input */
  float* grad_src_v;         /* DEVICE shape of src_v; overwritten                     */
} gbw_acoustic3d_adjoint;

/* Bytes of device workspace for gbw_acoustic3d_forward (pass 0),
 * gbw_acoustic3d_backward without (pass 1) or with (pass 2) the density
 * gradient; 0 if the problem is invalid. */
size_t gbw_acoustic3d_workspace_bytes(const gbw_acoustic3d_problem* prob, int pass);

/* Bytes of the checkpoint buffer (0 when checkpoint_every == 0). */
size_t gbw_acoustic3d_checkpoint_bytes(const gbw_acoustic3d_problem* prob);

/* Forward modelling: all nt steps, seismograms, energy maps and, when
 * checkpoint_every > 0, the checkpoints the gradient needs.
 * Replaces acoustic_forward_kernel_3d (acoustic_3d.py:275-476). */
int gbw_acoustic3d_forward(const gbw_acoustic3d_problem* prob, const gbw_acoustic3d_inputs* in,
                           const gbw_acoustic3d_outputs* out, void* workspace,
                           size_t workspace_bytes, void* stream /* hipStream_t */);

/* Gradient of sum(g_p*p + g_vx*vx + g_vy*vy + g_vz*vz) w.r.t. vp (through
 * alpha1 and pad_model_3d), rho and src_v: the hand-written adjoint of the 3D
 * step, recomputing each checkpoint interval.  Replaces PyTorch autograd
 * through torch.utils.checkpoint(acoustic_step_forward_3d) (acoustic_3d.py:415-437). */
int gbw_acoustic3d_backward(const gbw_acoustic3d_problem* prob, const gbw_acoustic3d_inputs* in,
                            const gbw_acoustic3d_adjoint* adj, void* workspace,
                            size_t workspace_bytes, void* stream /* hipStream_t */);

#ifdef __cplusplus
}
#endif

#endif /* GBW_ACOUSTIC3D_H */
