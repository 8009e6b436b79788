/*
 * gbw_acoustic2d.h -- C ABI of the MI355X (gfx950) 2D acoustic propagator.
 *
 * The library replaces the time-stepping operator of GeoBrain's 2D acoustic
 * propagator:
 *   acoustic_forward_kernel(nx, nz, dx, dz, nt, dt, nabc, free_surface,
 *                           src_x, src_z, src_n, src_v, rcv_x, rcv_z, rcv_n,
 *                           damp, vp, rho, checkpoint_segments, device, dtype,
 *                           snapshot_every)
 *       geobrain/physics/wave/kernels/acoustic_2d.py:294-501
 * together with the per-step body it drives,
 *   acoustic_step_forward(...)   geobrain/physics/wave/kernels/acoustic_2d.py:123-291
 * and the gradient PyTorch autograd computes for it through
 * torch.utils.checkpoint (acoustic_2d.py:448-460).  The reference is pure
 * Python and binds no C interface; these entry points are what its Python
 * layer would bind with ctypes (see INTEGRATION.md for the stub).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative gbw_status; the
 *    message of the last failure on the calling thread is gbw_last_error().
 *  - Arrays marked DEVICE are device pointers (e.g. torch tensor data_ptr()),
 *    contiguous, row-major, float32 unless stated.  Arrays marked HOST are host
 *    pointers read during the call only.
 *  - All device work is enqueued on `stream`; nothing is allocated: the caller
 *    passes a workspace of at least *_workspace_bytes() bytes and, for the
 *    gradient, a checkpoint buffer of gbw_acoustic2d_checkpoint_bytes() bytes.
 *  - Inputs are never modified.
 */
#ifndef GBW_ACOUSTIC2D_H
#define GBW_ACOUSTIC2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBW_ABI_VERSION 6

typedef enum {
  GBW_OK = 0,
  GBW_ERR_INVALID = -1,     /* bad sizes / flags (reference: ValueError)              */
  GBW_ERR_INDEX = -2,       /* source or receiver index out of range (IndexError)     */
  GBW_ERR_WORKSPACE = -3,   /* workspace or checkpoint buffer too small               */
  GBW_ERR_HIP = -4,         /* a HIP runtime call failed                              */
  GBW_ERR_UNSUPPORTED = -5  /* valid request the library does not implement           */
} gbw_status;

/* Scalars of acoustic_forward_kernel (acoustic_2d.py:294-317) plus launch
 * choices.  dx is accepted for signature parity; the scheme uses dz on both
 * axes, as the reference does (acoustic_2d.py:407,410). */
typedef struct gbw_acoustic2d_problem {
  int32_t nx, nz;            /* model grid (unpadded)                                  */
  float dx, dz;              /* grid spacing (m)                                       */
  int32_t nt;                /* time steps                                             */
  float dt;                  /* time step (s)                                          */
  int32_t nabc;              /* absorbing layers per side (>= 1)                       */
  int32_t free_surface;      /* 0 / 1                                                  */
  int32_t src_n;             /* shots                                                  */
  int32_t rcv_n;             /* receivers shared by every shot                         */
  int32_t wavelet_shared;    /* 1: src_v is (nt,) and feeds every shot; 0: (src_n, nt) */
                             /* (every point of a shot); 2: (src_n, src_points, nt)     */
  int32_t energy_chunk;      /* steps per chunk whose energy sums are added (0 = nt)   */
  int32_t snapshot_every;    /* >0: copy shot-0 pressure after every chunk of this size */
  int32_t checkpoint_every;  /* k>0: keep the state every k steps for the gradient     */
  int32_t checkpoint_bf16;   /* 1: store checkpoints as bf16 with per-field fp32 scale */
  int32_t tile_rows;         /* launch choice: rows per tile: 16 or 32 for one-step     */
                             /* launches; 24 selects 24 x 96 tiles for two-step ones    */
                             /* (else 16 x 128); 0 = default                            */
  int32_t shots_per_block;   /* launch choice: shots one workgroup steps (0 = default); */
                             /* the z-marching kernels take at most 8 of them, the      */
                             /* gradient's tiled recompute beside them all              */
  int32_t march;             /* launch choice for the gradient: 0 auto, 1 tiles,        */
                             /* 2 z-marching recompute strips (one step per launch),    */
                             /* 3 z-marching two-step adjoint (with a rho gradient the  */
                             /* tiled adjoint runs); auto takes 3 for 8 or more shots   */
  int32_t steps_per_launch;  /* launch choice: time steps per forward/recompute launch */
                             /* (1 or 2; 0 = default 2; two-step launches tile 16 rows) */
  int32_t fwd_march;         /* launch choice for two-step forward and recompute        */
                             /* launches: 1 tiles in shot groups; 2 z-marching          */
                             /* wavefronts (without a free surface; the rho gradient's  */
                             /* recompute keeps tiles); 3 two-phase tiles (the first    */
                             /* multiple of 256 tiles in shot groups, the rest in       */
                             /* groups of 8 shots); 0 auto: 3 where the grid has more   */
                             /* than 256 two-step tiles and groups of more than 8       */
                             /* shots, else 1                                           */
  int32_t src_points;        /* point sources per shot (0 or 1: one; K > 1: src_x and   */
                             /* src_z are (src_n, K), distinct cells within a shot, as  */
                             /* the reference's 2-D positions, acoustic_2d.py:234-237;  */
                             /* such problems run one-step tiled launches, and the      */
                             /* batched calls refuse them)                              */
} gbw_acoustic2d_problem;

/* Geometry and model (acoustic_2d.py:303-312).  Positions are unpadded model
 * indices, as the reference's kernel receives them.  With `geometry` set (a
 * buffer prepared by gbw_acoustic2d_geometry_setup), the positions are not
 * read: the call uses the prepared source cells and receiver maps and does no
 * host work, copy or synchronisation, so it can be captured in a HIP graph. */
typedef struct gbw_acoustic2d_inputs {
  const int64_t* src_x;      /* HOST (src_n,), or (src_n, src_points) row-major        */
  const int64_t* src_z;      /* HOST (src_n,), or (src_n, src_points) row-major        */
  const int64_t* rcv_x;      /* HOST (rcv_n,)                                          */
  const int64_t* rcv_z;      /* HOST (rcv_n,)                                          */
  const float* src_v;        /* DEVICE (src_n, nt), (nt,) or (src_n, src_points, nt)   */
  const float* damp;         /* DEVICE (nz + 2 nabc, nx + 2 nabc)                      */
  const float* vp;           /* DEVICE (nz, nx)                                        */
  const float* rho;          /* DEVICE (nz, nx)                                        */
  const void* geometry;      /* DEVICE prepared geometry buffer, or NULL               */
} gbw_acoustic2d_inputs;

/* Outputs of the forward pass (the dict of acoustic_2d.py:491-500).  Any
 * pointer may be NULL to skip that output. */
typedef struct gbw_acoustic2d_outputs {
  float* rec_p;              /* DEVICE (src_n, nt, rcv_n)  key 'p'                     */
  float* rec_vx;             /* DEVICE (src_n, nt, rcv_n)  key 'vx'                    */
  float* rec_vz;             /* DEVICE (src_n, nt, rcv_n)  key 'vz'                    */
  float* energy_p;           /* DEVICE (nz, nx)  key 'forward_wavefield_p'             */
  float* energy_vx;          /* DEVICE (nz, nx)  key 'forward_wavefield_vx'            */
  float* energy_vz;          /* DEVICE (nz, nx)  key 'forward_wavefield_vz'            */
  float* snapshots;          /* DEVICE (ceil(nt/snapshot_every), nz, nx)  'snapshots'  */
  void* checkpoints;         /* DEVICE checkpoint_bytes(), needed when checkpoint_every */
} gbw_acoustic2d_outputs;

/* Gradient request: adjoint sources are dLoss/d(seismogram).  Any g_* may be
 * NULL (zero); any grad_* may be NULL (not wanted). */
typedef struct gbw_acoustic2d_adjoint {
  const float* g_p;          /* DEVICE (src_n, nt, rcv_n)                              */
  const float* g_vx;         /* DEVICE (src_n, nt, rcv_n)                              */
  const float* g_vz;         /* DEVICE (src_n, nt, rcv_n)                              */
  const void* checkpoints;   /* DEVICE buffer filled by gbw_acoustic2d_forward         */
  float* grad_vp;            /* DEVICE (nz, nx)   overwritten                          */
  float* grad_rho;           /* DEVICE (nz, nx)   overwritten; density as an input     */
  float* grad_src_v;         /* DEVICE shape of src_v; overwritten                     */
} gbw_acoustic2d_adjoint;

int gbw_abi_version(void);
const char* gbw_last_error(void);

/* Bytes of device workspace for gbw_acoustic2d_forward (pass 0) or
 * gbw_acoustic2d_backward (pass 1); 0 if the problem is invalid. */
size_t gbw_acoustic2d_workspace_bytes(const gbw_acoustic2d_problem* prob, int pass);

/* Bytes of the checkpoint buffer (0 when checkpoint_every == 0). */
size_t gbw_acoustic2d_checkpoint_bytes(const gbw_acoustic2d_problem* prob);

/* Forward modelling: all nt steps, seismograms, energy maps, snapshots and,
 * when checkpoint_every > 0, the checkpoints the gradient needs.
 * Replaces acoustic_forward_kernel (acoustic_2d.py:294-501). */
int gbw_acoustic2d_forward(const gbw_acoustic2d_problem* prob,
                           const gbw_acoustic2d_inputs* in,
                           const gbw_acoustic2d_outputs* out,
                           void* workspace, size_t workspace_bytes,
                           void* stream /* hipStream_t */);

/* Gradient of sum(g_p*p + g_vx*vx + g_vz*vz) w.r.t. vp (through alpha1 and
 * pad_model), rho and src_v: the hand-written adjoint of the forward step,
 * recomputing each checkpoint interval.  Replaces PyTorch autograd through
 * torch.utils.checkpoint(acoustic_step_forward) (acoustic_2d.py:448-460). */
int gbw_acoustic2d_backward(const gbw_acoustic2d_problem* prob,
                            const gbw_acoustic2d_inputs* in,
                            const gbw_acoustic2d_adjoint* adj,
                            void* workspace, size_t workspace_bytes,
                            void* stream /* hipStream_t */);

/* Source cells and receiver maps of a problem, prepared once for calls that
 * repeat the same geometry (an FWI loop, graph capture): the part of
 * acoustic_forward_kernel that indexes with the positions (acoustic_2d.py:
 * 382-385, 228-237, 271-274; IndexError semantics as there).
 * gbw_acoustic2d_geometry_bytes gives the buffer size (0 if the problem is
 * invalid); gbw_acoustic2d_geometry_setup validates the positions in `in`,
 * writes the buffer (DEVICE) and synchronises `stream`.  Calls then pass it as
 * inputs.geometry; they check that their problem has the grid, absorbing
 * width, free surface, shot and receiver counts it was prepared for.  The
 * buffer must stay unchanged while calls use it; preparing it again (for new
 * positions) replaces it. */
size_t gbw_acoustic2d_geometry_bytes(const gbw_acoustic2d_problem* prob);
int gbw_acoustic2d_geometry_setup(const gbw_acoustic2d_problem* prob,
                                  const gbw_acoustic2d_inputs* in, void* geometry,
                                  size_t geometry_bytes, void* stream /* hipStream_t */);

/* Size of the device array behind gbw_acoustic2d_forward_noalloc (per GPU). */
#define GBW_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* Forward modelling that allocates nothing.  No counterpart in the reference:
 * it is gbw_acoustic2d_forward with the workspace taken from a fixed device
 * array of GBW_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it
 * runs on until the process ends, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The caller still passes inputs and outputs (and the
 * checkpoint buffer when checkpoint_every > 0).  A problem whose forward
 * workspace (gbw_acoustic2d_workspace_bytes(prob, 0)) exceeds the array is
 * rejected with GBW_ERR_WORKSPACE before anything is enqueued.  The call makes
 * no hipMalloc / hipFree and no other allocation on the device, and never
 * records timing events.  Every call on one GPU uses that GPU's one array:
 * calls that overlap on one GPU (two host threads, or two streams whose work
 * overlaps) race; calls in order on one stream do not.  Results equal
 * gbw_acoustic2d_forward's bit for bit. */
int gbw_acoustic2d_forward_noalloc(const gbw_acoustic2d_problem* prob,
                                   const gbw_acoustic2d_inputs* in,
                                   const gbw_acoustic2d_outputs* out,
                                   void* stream /* hipStream_t */);

/* Batched forward modelling of n independent problems (uneven grids, step
 * counts and shot counts) as one launch sequence: one launch per two time steps
 * covers every problem still running.  No counterpart in the reference, which
 * models one problem per acoustic_forward_kernel call (acoustic_2d.py:294-501);
 * this runs n such calls' forward passes together.  Each problem is described
 * exactly as for gbw_acoustic2d_forward (seismograms only: energy maps,
 * snapshots and checkpoints must be NULL / 0; steps_per_launch must agree).
 *  - gbw_acoustic2d_batch_setup validates the geometry (IndexError semantics),
 *    builds coefficients and receiver maps from the inputs and stores the
 *    output pointers in the workspace; it synchronises `stream`.  Call it
 *    again after changing any input.
 *  - gbw_acoustic2d_batch_forward only enqueues kernels (no allocation, copy
 *    or synchronisation), so it can be captured in a HIP graph and replayed.
 *    With done_ticks (DEVICE, n + 1 entries) it writes the device wall clock
 *    (gbw_device_wall_clock_khz) at its start into done_ticks[0] and when
 *    problem i finishes into done_ticks[1 + i], where i is its position in the
 *    problem list sorted by nt, longest first (ties keep the caller's order). */
size_t gbw_acoustic2d_batch_workspace_bytes(const gbw_acoustic2d_problem* probs, int n);
int gbw_acoustic2d_batch_setup(const gbw_acoustic2d_problem* probs,
                               const gbw_acoustic2d_inputs* ins,
                               const gbw_acoustic2d_outputs* outs, int n, void* workspace,
                               size_t workspace_bytes, void* stream /* hipStream_t */);
int gbw_acoustic2d_batch_forward(const gbw_acoustic2d_problem* probs, int n, void* workspace,
                                 size_t workspace_bytes, unsigned long long* done_ticks,
                                 void* stream /* hipStream_t */);
int gbw_device_wall_clock_khz(void);

/* Batched gradient of n independent problems (uneven grids, step counts and
 * shot counts): the forward with checkpoints, then the recompute and adjoint,
 * of all n problems as one launch sequence each.  No single counterpart in the
 * reference, which differentiates one acoustic_forward_kernel call at a time
 * with PyTorch autograd through torch.utils.checkpoint
 * (acoustic_2d.py:448-460); this runs n such gradients together.  Problems are
 * aligned at their last time step: every problem's checkpoint intervals then
 * coincide (checkpoint_every, equal for all problems and even, 0 = automatic),
 * and each launch advances every problem that has started by two steps.
 * Each problem's results equal its own gbw_acoustic2d_forward /
 * gbw_acoustic2d_backward up to the order of FP32 sums (bit for bit for the
 * seismograms).
 *  - gbw_acoustic2d_batch_gradient_setup validates the geometry (IndexError
 *    semantics), builds coefficients and receiver maps and stores the output
 *    and adjoint pointers: outs[i] takes seismograms only (energy maps and
 *    snapshots NULL); adjs[i] gives the adjoint sources g_p / g_vx / g_vz (any
 *    may be NULL) and the outputs grad_vp and grad_src_v (either may be NULL;
 *    grad_rho must be NULL; checkpoints is ignored: they live in the
 *    workspace).  It synchronises `stream`.  Call it again after changing any
 *    input or pointer.
 *  - gbw_acoustic2d_batch_gradient_forward runs the forward pass (seismograms
 *    into outs, checkpoints into the workspace); _backward reads the adjoint
 *    sources and writes the gradients.  Both only enqueue kernels (no
 *    allocation, copy or synchronisation), so forward, the caller's misfit and
 *    backward can be captured in one HIP graph.  With done_ticks (DEVICE,
 *    n + 1 entries) forward writes the device wall clock at its start into
 *    done_ticks[0] and backward writes it into done_ticks[1 + i] when problem i's
 *    gradient is complete, i its position in the list sorted by nt, longest
 *    first (ties keep the caller's order). */
size_t gbw_acoustic2d_batch_gradient_workspace_bytes(const gbw_acoustic2d_problem* probs, int n);
int gbw_acoustic2d_batch_gradient_setup(const gbw_acoustic2d_problem* probs,
                                        const gbw_acoustic2d_inputs* ins,
                                        const gbw_acoustic2d_outputs* outs,
                                        const gbw_acoustic2d_adjoint* adjs, int n,
                                        void* workspace, size_t workspace_bytes,
                                        void* stream /* hipStream_t */);
int gbw_acoustic2d_batch_gradient_forward(const gbw_acoustic2d_problem* probs, int n,
                                          void* workspace, size_t workspace_bytes,
                                          unsigned long long* done_ticks,
                                          void* stream /* hipStream_t */);
int gbw_acoustic2d_batch_gradient_backward(const gbw_acoustic2d_problem* probs, int n,
                                           void* workspace, size_t workspace_bytes,
                                           unsigned long long* done_ticks,
                                           void* stream /* hipStream_t */);

/* Measurement hook, no counterpart in the reference (for benchmarks and
 * profilers that need per-kernel times from a live run).  While enabled, each
 * time-step kernel launch is bracketed by two events on the caller's stream
 * and every forward / backward call waits for its stream before returning.
 * gbw_set_kernel_timing(1) enables it and clears the totals, (0) disables it.
 * gbw_kernel_times writes one line per kernel, "name launches total_ms", into
 * buf (NUL-terminated, truncated to len) and returns the bytes it needs. */
void gbw_set_kernel_timing(int on);
size_t gbw_kernel_times(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* GBW_ACOUSTIC2D_H */
