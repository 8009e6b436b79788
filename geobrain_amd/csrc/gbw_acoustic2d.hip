// MI355X (gfx950) implementation of the 2D acoustic staggered-grid propagator
// behind the C ABI declared in include/gbw_acoustic2d.h.
//
// Scheme (geobrain/physics/wave/kernels/acoustic_2d.py:211-285): per step
//   p  <- (1-k1) p  - a1 * D(vx, vz)        on rows [f+1, NZ-3], cols [2, NX-3]
//   p[src] += dt * w                        (every shot, its own cell)
//   p[f-1, :] = -p[f+1, :]                  (free surface only)
//   vx <- (1-k2) vx - a2 * Gx(p)            on rows [f, NZ-2],  cols [1, NX-3]
//   vz <- (1-k3) vz - a2 * Gz(p)            on rows [f, NZ-3],  cols [1, NX-2]
//   vz[f-1, :] = vz[f, :]                   (free surface only)
//   record p, vx, vz at the receiver cells; energy += sum_s field^2 (interior)
// with D, Gx, Gz the 4th-order staggered differences (c1 = 9/8, c2 = -1/24).
// Every floating-point operation is written in the reference's order and the
// library is built with -ffp-contract=off, so each result is rounded exactly
// as the reference's FP32 PyTorch ops round it.
//
// Data layout in HBM (one problem):
//   plane   = NZ rows x PX floats, PX = NX rounded up to 64 (256-byte rows)
//   state   = [field P|VX|VZ][shot][plane]   (ping-pong pair A/B)
//   coefs   = [omk1|a1|omk2|omk3|a2][plane]  (1-kappa and alpha, built on device)
//   checkpoints: [slot][field][shot][plane]; slot c = state at step c*k.
//
// Kernels
//   fwd_step   one time step for one TZ x 64 tile and a group of shots.  The
//              tile's vx/vz halo is staged in LDS, p' is computed on the tile
//              plus a 1/2-cell halo (recomputed by neighbours instead of a
//              second launch), then vx'/vz' on the tile.  Fields in: 12 B per
//              cell, out: 12 B per cell.
//   adj_step   the exact adjoint of fwd_step (same footprint, mirrored), with
//              the zero-lag imaging condition d alpha1 += -lambda_p' * D^n read
//              from the bracket stored during the checkpoint-interval recompute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "gbw_acoustic2d.h"

namespace {

constexpr float kC1 = 9.0f / 8.0f;
constexpr float kC2 = (float)(-1.0 / 24.0);
constexpr int TX = 64;        // tile width: one 64-lane wavefront per row
constexpr int NTHREADS = 256; // four wavefronts per workgroup

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(GBW_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_CHECK()                                                                 \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess)                                                              \
      return fail(GBW_ERR_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e_)); \
  } while (0)

struct Dims {
  int nz, nx, nabc, NZ, NX, PX, f, fs;
  long long plane;  // NZ * PX
};

// ---------------------------------------------------------------------------
// Coefficients (acoustic_2d.py:374-375 pad_model, :402-420 precompute)
// ---------------------------------------------------------------------------
__global__ void coef_kernel(Dims d, const float* __restrict__ vp, const float* __restrict__ rho,
                            const float* __restrict__ damp, float dt, float dz,
                            float* __restrict__ coef) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= d.PX || i >= d.NZ) return;
  const long long q = (long long)i * d.PX + j;
  float omk1 = 1.f, a1 = 0.f, omk2 = 1.f, omk3 = 1.f, a2 = 0.f;
  if (j < d.NX) {
    const int a = min(max(i - d.nabc, 0), d.nz - 1);
    const int b = min(max(j - d.nabc, 0), d.nx - 1);
    const float c = vp[(long long)a * d.nx + b];
    const float den = rho[(long long)a * d.nx + b];
    float t = den * c;
    t = t * c;
    t = t * dt;
    a1 = t / dz;
    const float* drow = damp + (long long)i * d.NX;
    const float k1 = drow[j] * dt;
    omk1 = 1.f - k1;
    const float dd = den * dz;
    const float r = 1.f / dd;
    a2 = r * dt;
    float k2 = 0.f;
    if (j >= 1 && j <= d.NX - 3) {
      float s = drow[j] + drow[j + 1];
      s = 0.5f * s;
      k2 = s * dt;
    }
    omk2 = 1.f - k2;
    float k3 = 0.f;
    if (i >= d.f && i <= d.NZ - 3) {
      float s = drow[j] + drow[j + d.NX];
      s = 0.5f * s;
      k3 = s * dt;
    }
    omk3 = 1.f - k3;
  }
  coef[q] = omk1;
  coef[d.plane + q] = a1;
  coef[2 * d.plane + q] = omk2;
  coef[3 * d.plane + q] = omk3;
  coef[4 * d.plane + q] = a2;
}

__global__ void fill_kernel(float* __restrict__ p, long long n, float v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void fill_int_kernel(int* __restrict__ p, long long n, int v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// receiver map: rhead[cell] = first receiver recorded at that cell; lists are
// ordered by receiver index (host-built rnext), so duplicate receivers sum in
// index order.
__global__ void rmap_scatter_kernel(int* __restrict__ rhead, const int* __restrict__ heads_cell,
                                    const int* __restrict__ heads_r, int n_heads) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_heads) rhead[heads_cell[i]] = heads_r[i];
}

// ---------------------------------------------------------------------------
// Forward step
// ---------------------------------------------------------------------------
struct FwdArgs {
  Dims d;
  int S, spb, tiles_x;           // shots, shots per block, tiles along x
  const float* P0; const float* VX0; const float* VZ0;  // input state fields (shot-major)
  float* P1; float* VX1; float* VZ1;                    // output state fields
  const float* coef;
  const int* src_cell;           // [S] padded cell index (row * PX + col)
  const float* wav; long long wav_stride;  // w[s * wav_stride + it]
  int it, nt; float dt;
  const int* rhead; const int* rnext; const unsigned char* rrow; int R;
  float* rec_p; float* rec_vx; float* rec_vz;  // [S][nt][R] or null
  float* energy;                 // [group][3][nz][nx] or null
  float* dbuf;                   // [S][plane] bracket D of this step, or null
  float* gxbuf; float* gzbuf;    // [S][plane] vx / vz brackets of this step, or null
};

// RECOMP = false: forward pass (receivers, energy); true: checkpoint-interval
// recompute for the adjoint (stores the brackets the adjoint needs instead).
template <int TZ, bool RECOMP>
__global__ __launch_bounds__(NTHREADS) void fwd_step_kernel(FwdArgs a) {
  __shared__ float sVX[TZ + 3][TX + 6];  // rows z0-1..z0+TZ+1, cols x0-3..x0+TX+2
  __shared__ float sVZ[TZ + 6][TX + 3];  // rows z0-3..z0+TZ+2, cols x0-1..x0+TX+1
  __shared__ float sP[TZ + 3][TX + 3];   // rows z0-1..z0+TZ+1, cols x0-1..x0+TX+1
  constexpr int RPT = TZ / 4;            // tile rows per wavefront
  const Dims d = a.d;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile = blockIdx.x;
  const int x0 = (tile % a.tiles_x) * TX;
  const int z0 = (tile / a.tiles_x) * TZ;
  const int s_begin = blockIdx.y * a.spb;
  const int s_end = min(a.S, s_begin + a.spb);
  const float* omk1 = a.coef;
  const float* a1 = a.coef + d.plane;
  const float* omk2 = a.coef + 2 * d.plane;
  const float* omk3 = a.coef + 3 * d.plane;
  const float* a2 = a.coef + 4 * d.plane;
  const long long fstride = (long long)a.S * d.plane;  // between P, VX, VZ of one state

  float e_p[RPT], e_vx[RPT], e_vz[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) e_p[k] = e_vx[k] = e_vz[k] = 0.f;

  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const float* P0 = a.P0 + so;
    const float* VX0 = a.VX0 + so;
    const float* VZ0 = a.VZ0 + so;
    for (int r = wave; r < TZ + 3; r += 4) {
      const int gz = z0 - 1 + r;
      const bool zin = gz >= 0 && gz < d.NZ;
      for (int c = lane; c < TX + 6; c += 64) {
        const int gx = x0 - 3 + c;
        sVX[r][c] = (zin && gx >= 0 && gx < d.NX) ? VX0[(long long)gz * d.PX + gx] : 0.f;
      }
    }
    for (int r = wave; r < TZ + 6; r += 4) {
      const int gz = z0 - 3 + r;
      const bool zin = gz >= 0 && gz < d.NZ;
      for (int c = lane; c < TX + 3; c += 64) {
        const int gx = x0 - 1 + c;
        sVZ[r][c] = (zin && gx >= 0 && gx < d.NX) ? VZ0[(long long)gz * d.PX + gx] : 0.f;
      }
    }
    __syncthreads();
    const int scell = a.src_cell[s];
    const float supd = a.dt * a.wav[(long long)s * a.wav_stride + a.it];
    // p' on the tile and its halo
    for (int r = wave; r < TZ + 3; r += 4) {
      const int gz = z0 - 1 + r;
      const bool zin = gz >= 0 && gz < d.NZ;
      const bool zreg = gz >= d.f + 1 && gz <= d.NZ - 3;
      for (int c = lane; c < TX + 3; c += 64) {
        const int gx = x0 - 1 + c;
        float pv = 0.f;
        if (zin && gx >= 0 && gx < d.NX) {
          const long long q = (long long)gz * d.PX + gx;
          pv = P0[q];
          if (zreg && gx >= 2 && gx <= d.NX - 3) {
            float s1 = sVX[r][c + 2] - sVX[r][c + 1];
            s1 = s1 + sVZ[r + 2][c];
            s1 = s1 - sVZ[r + 1][c];
            float s2 = sVX[r][c + 3] - sVX[r][c];
            s2 = s2 + sVZ[r + 3][c];
            s2 = s2 - sVZ[r][c];
            const float t1 = kC1 * s1;
            const float t2 = kC2 * s2;
            const float br = t1 + t2;
            const float u = omk1[q] * pv;
            pv = u - a1[q] * br;
            if constexpr (RECOMP) {
              if (r >= 1 && r <= TZ && c >= 1 && c <= TX) a.dbuf[so + q] = br;
            }
          }
          if ((int)q == scell) pv = pv + supd;
        }
        sP[r][c] = pv;
      }
    }
    __syncthreads();
    if (d.fs) {
      const int rm = d.f - 1 - (z0 - 1);  // rect row of f-1; f+1 is rect row rm+2
      if (rm >= 0 && rm <= TZ) {
        for (int c = tid; c < TX + 3; c += NTHREADS) sP[rm][c] = -sP[rm + 2][c];
      }
      __syncthreads();
    }
    // v' on the tile
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int tr = wave + 4 * k;
      const int gz = z0 + tr, gx = x0 + lane;
      if (gz < d.NZ && gx < d.NX) {
        const long long q = (long long)gz * d.PX + gx;
        const float pc = sP[tr + 1][lane + 1];
        float vx = sVX[tr + 1][lane + 3];
        float vz = sVZ[tr + 3][lane + 1];
        if (gz >= d.f && gz <= d.NZ - 2 && gx >= 1 && gx <= d.NX - 3) {
          const float d1 = sP[tr + 1][lane + 2] - pc;
          const float d2 = sP[tr + 1][lane + 3] - sP[tr + 1][lane];
          const float br = kC1 * d1 + kC2 * d2;
          const float u = omk2[q] * vx;
          vx = u - a2[q] * br;
          if constexpr (RECOMP) {
            if (a.gxbuf) a.gxbuf[so + q] = br;
          }
        }
        if (gz >= d.f && gz <= d.NZ - 3 && gx >= 1 && gx <= d.NX - 2) {
          const float d1 = sP[tr + 2][lane + 1] - pc;
          const float d2 = sP[tr + 3][lane + 1] - sP[tr][lane + 1];
          const float br = kC1 * d1 + kC2 * d2;
          const float u = omk3[q] * vz;
          vz = u - a2[q] * br;
          if constexpr (RECOMP) {
            if (a.gzbuf) a.gzbuf[so + q] = br;
          }
        }
        a.P1[so + q] = pc;
        a.VX1[so + q] = vx;
        if (!(d.fs && gz == d.f - 1)) a.VZ1[so + q] = vz;
        if (d.fs && gz == d.f) a.VZ1[so + q - d.PX] = vz;  // vz[f-1] = vz[f]
        if constexpr (!RECOMP) {
          if (a.rec_p && a.rrow[gz]) {
            for (int h = a.rhead[q]; h >= 0; h = a.rnext[h]) {
              const long long o = ((long long)s * a.nt + a.it) * a.R + h;
              a.rec_p[o] = pc;
              a.rec_vx[o] = vx;
              a.rec_vz[o] = vz;
            }
          }
          e_p[k] = e_p[k] + pc * pc;
          e_vx[k] = e_vx[k] + vx * vx;
          e_vz[k] = e_vz[k] + vz * vz;
        }
      }
    }
    __syncthreads();
  }
  (void)fstride;
  if (!RECOMP && a.energy && s_begin < s_end) {
    const long long inner = (long long)d.nz * d.nx;
    float* eg = a.energy + (long long)blockIdx.y * 3 * inner;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int ia = z0 + wave + 4 * k - d.nabc;
      const int jb = x0 + lane - d.nabc;
      if (ia >= 0 && ia < d.nz && jb >= 0 && jb < d.nx) {
        const long long o = (long long)ia * d.nx + jb;
        eg[o] = eg[o] + e_p[k];
        eg[inner + o] = eg[inner + o] + e_vx[k];
        eg[2 * inner + o] = eg[2 * inner + o] + e_vz[k];
      }
    }
  }
}

// totals += sum_g partial[g]; partial = 0   (acoustic_2d.py:479-481)
__global__ void energy_flush_kernel(float* __restrict__ part, int groups, long long inner,
                                    float* __restrict__ ep, float* __restrict__ evx,
                                    float* __restrict__ evz) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * inner) return;
  const int fld = (int)(i / inner);
  const long long o = i - (long long)fld * inner;
  float s = part[i];
  part[i] = 0.f;
  for (int g = 1; g < groups; ++g) {
    float* pg = part + (long long)g * 3 * inner + i;
    s = s + *pg;
    *pg = 0.f;
  }
  float* dst = fld == 0 ? ep : (fld == 1 ? evx : evz);
  if (dst) dst[o] = dst[o] + s;
}

// snapshot of shot-0 pressure interior (acoustic_2d.py:483-487)
__global__ void snapshot_kernel(Dims d, const float* __restrict__ P, float* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b < d.nx && a < d.nz) out[(long long)a * d.nx + b] = P[(long long)(a + d.nabc) * d.PX + b + d.nabc];
}

// ---------------------------------------------------------------------------
// Adjoint step
// ---------------------------------------------------------------------------
struct AdjArgs {
  Dims d;
  int S, spb, tiles_x;
  const float* LP0; const float* LX0; const float* LZ0;  // adjoint of state it+1 (records of it injected)
  float* LP1; float* LX1; float* LZ1;                    // adjoint of state it (records of it-1 injected)
  const float* coef;
  const float* dbuf;                       // [S][plane] bracket D^it
  const float* gxbuf; const float* gzbuf;  // [S][plane] vx/vz brackets of step it (rho gradient) or null
  const int* src_cell;
  int it, nt; float dt;
  const int* rhead; const int* rnext; const unsigned char* rrow; int R;
  const float* g_p; const float* g_vx; const float* g_vz;  // [S][nt][R] or null
  float* gA1;   // [group][plane]
  float* gA2;   // [group][plane] or null
  float* gw;    // [S][nt] or null
};

template <int TZ>
__global__ __launch_bounds__(NTHREADS) void adj_step_kernel(AdjArgs a) {
  __shared__ float sQX[TZ + 3][TX + 6];  // a2*lambda_vx on the vx region, rows z0-1.., cols x0-3..
  __shared__ float sQZ[TZ + 6][TX + 3];  // a2*lambda_vz on the vz region, rows z0-3.., cols x0-1..
  __shared__ float sLP[TZ + 3][TX + 3];  // lambda_p' rows z0-1.., cols x0-1..
  __shared__ float sR[TZ + 3][TX + 3];   // a1*lambda_p' on the p region
  constexpr int RPT = TZ / 4;
  const Dims d = a.d;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tile = blockIdx.x;
  const int x0 = (tile % a.tiles_x) * TX;
  const int z0 = (tile / a.tiles_x) * TZ;
  const int s_begin = blockIdx.y * a.spb;
  const int s_end = min(a.S, s_begin + a.spb);
  const float* omk1 = a.coef;
  const float* a1 = a.coef + d.plane;
  const float* omk2 = a.coef + 2 * d.plane;
  const float* omk3 = a.coef + 3 * d.plane;
  const float* a2 = a.coef + 4 * d.plane;

  float gacc[RPT], gacc2[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) gacc[k] = gacc2[k] = 0.f;

  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const float* LP0 = a.LP0 + so;
    const float* LX0 = a.LX0 + so;
    const float* LZ0 = a.LZ0 + so;
    for (int r = wave; r < TZ + 3; r += 4) {
      const int gz = z0 - 1 + r;
      const bool zreg = gz >= d.f && gz <= d.NZ - 2;
      for (int c = lane; c < TX + 6; c += 64) {
        const int gx = x0 - 3 + c;
        float v = 0.f;
        if (zreg && gx >= 1 && gx <= d.NX - 3) {
          const long long q = (long long)gz * d.PX + gx;
          v = a2[q] * LX0[q];
        }
        sQX[r][c] = v;
      }
    }
    for (int r = wave; r < TZ + 6; r += 4) {
      const int gz = z0 - 3 + r;
      const bool zreg = gz >= d.f && gz <= d.NZ - 3;
      for (int c = lane; c < TX + 3; c += 64) {
        const int gx = x0 - 1 + c;
        float v = 0.f;
        if (zreg && gx >= 1 && gx <= d.NX - 2) {
          const long long q = (long long)gz * d.PX + gx;
          float lz = LZ0[q];
          if (d.fs && gz == d.f) lz += LZ0[q - d.PX];  // transpose of vz[f-1] = vz[f]
          v = a2[q] * lz;
        }
        sQZ[r][c] = v;
      }
    }
    __syncthreads();
    // lambda_p' on the tile and its halo
    for (int r = wave; r < TZ + 3; r += 4) {
      const int gz = z0 - 1 + r;
      const bool zin = gz >= 0 && gz < d.NZ;
      for (int c = lane; c < TX + 3; c += 64) {
        const int gx = x0 - 1 + c;
        float lp = 0.f, rr = 0.f;
        if (zin && gx >= 0 && gx < d.NX) {
          const long long q = (long long)gz * d.PX + gx;
          lp = LP0[q];
          const float az = kC1 * (sQZ[r + 2][c] - sQZ[r + 1][c]) + kC2 * (sQZ[r + 3][c] - sQZ[r][c]);
          const float ax = kC1 * (sQX[r][c + 2] - sQX[r][c + 1]) + kC2 * (sQX[r][c + 3] - sQX[r][c]);
          float acc = az;
          acc = acc + ax;
          lp = lp + acc;
          if (d.fs) {
            if (gz == d.f + 1) {
              // lambda_p'[f-1] = lambda_p[f-1] + c2 * q_z[f]: rows above f carry no q
              const float zero = 0.f;
              const float azm = kC1 * (zero - zero) + kC2 * (sQZ[r + 1][c] - zero);
              const float axm = kC1 * (zero - zero) + kC2 * (zero - zero);
              float accm = azm;
              accm = accm + axm;
              const float lpm = LP0[q - 2 * d.PX] + accm;
              lp = lp - lpm;
            } else if (gz == d.f - 1) {
              lp = 0.f;
            }
          }
          if (gz >= d.f + 1 && gz <= d.NZ - 3 && gx >= 2 && gx <= d.NX - 3) rr = a1[q] * lp;
        }
        sLP[r][c] = lp;
        sR[r][c] = rr;
      }
    }
    __syncthreads();
    const bool inject = a.it >= 1;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int tr = wave + 4 * k;
      const int gz = z0 + tr, gx = x0 + lane;
      if (gz < d.NZ && gx < d.NX) {
        const long long q = (long long)gz * d.PX + gx;
        const float lp = sLP[tr + 1][lane + 1];
        float lpo = lp;
        if (gz >= d.f + 1 && gz <= d.NZ - 3 && gx >= 2 && gx <= d.NX - 3) {
          gacc[k] = gacc[k] - lp * a.dbuf[so + q];
          lpo = omk1[q] * lp;
        }
        if ((int)q == a.src_cell[s] && a.gw) a.gw[(long long)s * a.nt + a.it] = a.dt * lp;
        float lx = LX0[q];
        float lz = LZ0[q];
        if (d.fs && gz == d.f) lz += LZ0[q - d.PX];
        if (d.fs && gz == d.f - 1) lz = 0.f;
        const bool xreg = gz >= d.f && gz <= d.NZ - 2 && gx >= 1 && gx <= d.NX - 3;
        const bool zreg = gz >= d.f && gz <= d.NZ - 3 && gx >= 1 && gx <= d.NX - 2;
        if (a.gA2) {
          if (zreg) gacc2[k] = gacc2[k] - lz * a.gzbuf[so + q];
          if (xreg) gacc2[k] = gacc2[k] - lx * a.gxbuf[so + q];
        }
        if (xreg) lx = omk2[q] * lx;
        if (zreg) lz = omk3[q] * lz;
        if (gx <= d.NX - 2)
          lx += kC1 * (sR[tr + 1][lane + 2] - sR[tr + 1][lane + 1]) +
                kC2 * (sR[tr + 1][lane + 3] - sR[tr + 1][lane]);
        if (gz <= d.NZ - 2)
          lz += kC1 * (sR[tr + 2][lane + 1] - sR[tr + 1][lane + 1]) +
                kC2 * (sR[tr + 3][lane + 1] - sR[tr][lane + 1]);
        if (inject && a.rrow[gz]) {
          for (int h = a.rhead[q]; h >= 0; h = a.rnext[h]) {
            const long long o = ((long long)s * a.nt + (a.it - 1)) * a.R + h;
            if (a.g_p) lpo += a.g_p[o];
            if (a.g_vx) lx += a.g_vx[o];
            if (a.g_vz) lz += a.g_vz[o];
          }
        }
        a.LP1[so + q] = lpo;
        a.LX1[so + q] = lx;
        a.LZ1[so + q] = lz;
      }
    }
    __syncthreads();
  }
  if (s_begin < s_end) {
    float* g1 = a.gA1 + (long long)blockIdx.y * d.plane;
    float* g2 = a.gA2 ? a.gA2 + (long long)blockIdx.y * d.plane : nullptr;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int gz = z0 + wave + 4 * k, gx = x0 + lane;
      if (gz < d.NZ && gx < d.NX) {
        const long long q = (long long)gz * d.PX + gx;
        g1[q] = g1[q] + gacc[k];
        if (g2) g2[q] = g2[q] + gacc2[k];
      }
    }
  }
}

// lambda^{nt} = scatter of the records of the last step (zeroed buffer in)
__global__ void adj_inject_kernel(Dims d, int S, int nt, int it, int R, const int* __restrict__ rcell,
                                  const int* __restrict__ rhead, const int* __restrict__ rnext,
                                  const float* __restrict__ g_p, const float* __restrict__ g_vx,
                                  const float* __restrict__ g_vz, float* LP, float* LX, float* LZ) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= S * R) return;
  const int s = idx / R, r = idx % R;
  const int cell = rcell[r];
  if (rhead[cell] != r) return;  // only the list head sums its cell
  float sp = 0.f, sx = 0.f, sz = 0.f;
  for (int h = r; h >= 0; h = rnext[h]) {
    const long long o = ((long long)s * nt + it) * R + h;
    if (g_p) sp += g_p[o];
    if (g_vx) sx += g_vx[o];
    if (g_vz) sz += g_vz[o];
  }
  const long long q = (long long)s * d.plane + cell;
  LP[q] = LP[q] + sp;
  LX[q] = LX[q] + sx;
  LZ[q] = LZ[q] + sz;
}

// chain rule alpha1 = den*c*c*dt/dz, alpha2 = dt/(den*dz) and the fold of
// pad_model's edge replication (acoustic_2d.py:407-410, :18-66)
__global__ void fold_kernel(Dims d, int groups, const float* __restrict__ gA1,
                            const float* __restrict__ gA2, const float* __restrict__ vp,
                            const float* __restrict__ rho, float dt, float dz,
                            float* __restrict__ grad_vp, float* __restrict__ grad_rho) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= d.nx || a >= d.nz) return;
  const int i0 = a == 0 ? 0 : a + d.nabc;
  const int i1 = a == d.nz - 1 ? d.NZ - 1 : a + d.nabc;
  const int j0 = b == 0 ? 0 : b + d.nabc;
  const int j1 = b == d.nx - 1 ? d.NX - 1 : b + d.nabc;
  const float c = vp[(long long)a * d.nx + b];
  const float den = rho[(long long)a * d.nx + b];
  float sv = 0.f, sr = 0.f;
  for (int i = i0; i <= i1; ++i)
    for (int j = j0; j <= j1; ++j) {
      const long long q = (long long)i * d.PX + j;
      float g1 = gA1[q];
      for (int g = 1; g < groups; ++g) g1 = g1 + gA1[(long long)g * d.plane + q];
      float fv = 2.f * den;
      fv = fv * c;
      fv = fv * dt;
      fv = fv / dz;
      sv += g1 * fv;
      if (grad_rho) {
        float g2 = 0.f;
        if (gA2) {
          g2 = gA2[q];
          for (int g = 1; g < groups; ++g) g2 = g2 + gA2[(long long)g * d.plane + q];
        }
        float f1 = c * c;
        f1 = f1 * dt;
        f1 = f1 / dz;
        float f2 = den * den;
        f2 = f2 * dz;
        f2 = dt / f2;
        sr += g1 * f1 - g2 * f2;
      }
    }
  grad_vp[(long long)a * d.nx + b] = sv;
  if (grad_rho) grad_rho[(long long)a * d.nx + b] = sr;
}

// wavelet gradient: per shot, or summed over shots for a shared wavelet
__global__ void wavelet_grad_kernel(const float* __restrict__ gw, int S, int nt, int shared,
                                    float* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (shared) {
    if (t >= nt) return;
    float s = 0.f;
    for (int k = 0; k < S; ++k) s += gw[(long long)k * nt + t];
    out[t] = s;
  } else {
    if (t >= S * nt) return;
    out[t] = gw[t];
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Plan {
  Dims d;
  int S, R, nt, k, nseg, TZ, spb, groups, tiles_x, tiles_z, echunk;
  size_t plane_bytes;
  // byte offsets into the workspace
  size_t off_coef, off_A, off_B, off_LA, off_LB, off_D, off_GX, off_GZ, off_E, off_G1, off_G2,
      off_GW, off_rhead, off_rnext, off_rrow, off_rcell, off_heads, off_src, total;
};

int make_plan(const gbw_acoustic2d_problem* p, int pass, Plan& pl) {
  if (!p) return fail(GBW_ERR_INVALID, "problem is NULL");
  if (p->nx <= 0 || p->nz <= 0)
    return fail(GBW_ERR_INVALID, "Grid dimensions must be positive: nx=" + std::to_string(p->nx) +
                                     ", nz=" + std::to_string(p->nz));
  if (p->nabc < 1) return fail(GBW_ERR_INVALID, "nabc must be >= 1");
  if (p->nt <= 0) return fail(GBW_ERR_INVALID, "nt must be positive");
  if (p->src_n <= 0) return fail(GBW_ERR_INVALID, "src_n must be positive");
  if (p->rcv_n < 0) return fail(GBW_ERR_INVALID, "rcv_n must be >= 0");
  if (!(p->dt > 0.f) || !(p->dz > 0.f)) return fail(GBW_ERR_INVALID, "dt and dz must be positive");
  if (p->checkpoint_bf16) return fail(GBW_ERR_UNSUPPORTED, "bf16 checkpoints are not built yet");
  Dims& d = pl.d;
  d.nz = p->nz; d.nx = p->nx; d.nabc = p->nabc;
  d.NZ = p->nz + 2 * p->nabc; d.NX = p->nx + 2 * p->nabc;
  d.PX = (d.NX + 63) / 64 * 64;
  d.fs = p->free_surface ? 1 : 0;
  d.f = d.fs ? p->nabc : 1;
  d.plane = (long long)d.NZ * d.PX;
  if (d.NZ < 6 || d.NX < 6) return fail(GBW_ERR_INVALID, "padded grid must be at least 6 x 6");
  if (d.plane >= (1LL << 31)) return fail(GBW_ERR_INVALID, "padded plane exceeds 2^31 cells");
  pl.S = p->src_n; pl.R = p->rcv_n; pl.nt = p->nt;
  pl.k = p->checkpoint_every > 0 ? std::min(p->checkpoint_every, p->nt) : 0;
  if (pass >= 1 && pl.k == 0) return fail(GBW_ERR_INVALID, "the gradient needs checkpoint_every > 0");
  pl.nseg = pl.k ? (p->nt + pl.k - 1) / pl.k : 0;
  pl.TZ = p->tile_rows == 16 ? 16 : 8;
  if (p->tile_rows != 0 && p->tile_rows != 8 && p->tile_rows != 16)
    return fail(GBW_ERR_INVALID, "tile_rows must be 0, 8 or 16");
  pl.tiles_x = (d.NX + TX - 1) / TX;
  pl.tiles_z = (d.NZ + pl.TZ - 1) / pl.TZ;
  const int tiles = pl.tiles_x * pl.tiles_z;
  int spb = p->shots_per_block;
  if (spb <= 0) {
    // enough workgroups to fill 256 CUs several times over, shots amortise coefficients
    const int want_groups = std::max(1, std::min(pl.S, (2048 + tiles - 1) / tiles));
    spb = (pl.S + want_groups - 1) / want_groups;
  }
  spb = std::min(spb, pl.S);
  pl.spb = spb;
  pl.groups = (pl.S + spb - 1) / spb;
  pl.echunk = p->energy_chunk > 0 ? std::min(p->energy_chunk, p->nt) : p->nt;
  pl.plane_bytes = (size_t)d.plane * sizeof(float);
  const size_t state = 3 * (size_t)pl.S * pl.plane_bytes;
  const size_t inner = (size_t)d.nz * d.nx * sizeof(float);
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += align256(bytes); return at; };
  pl.off_coef = take(5 * pl.plane_bytes);
  pl.off_A = take(state);
  pl.off_B = take(state);
  pl.off_LA = pl.off_LB = pl.off_D = pl.off_GX = pl.off_GZ = pl.off_G1 = pl.off_G2 = pl.off_GW = 0;
  pl.off_E = 0;
  if (pass == 0) {
    pl.off_E = take((size_t)pl.groups * 3 * inner);
  } else {
    pl.off_LA = take(state);
    pl.off_LB = take(state);
    pl.off_D = take((size_t)pl.k * pl.S * pl.plane_bytes);
    if (pass == 2) {
      pl.off_GX = take((size_t)pl.k * pl.S * pl.plane_bytes);
      pl.off_GZ = take((size_t)pl.k * pl.S * pl.plane_bytes);
      pl.off_G2 = take((size_t)pl.groups * pl.plane_bytes);
    }
    pl.off_G1 = take((size_t)pl.groups * pl.plane_bytes);
    pl.off_GW = take((size_t)pl.S * pl.nt * sizeof(float));
  }
  pl.off_rhead = take((size_t)d.plane * sizeof(int));
  pl.off_rnext = take((size_t)std::max(1, pl.R) * sizeof(int));
  pl.off_rcell = take((size_t)std::max(1, pl.R) * sizeof(int));
  pl.off_heads = take(2 * (size_t)std::max(1, pl.R) * sizeof(int));
  pl.off_rrow = take((size_t)d.NZ);
  pl.off_src = take((size_t)pl.S * sizeof(int));
  pl.total = o;
  return 0;
}

size_t checkpoint_bytes(const Plan& pl) {
  return pl.k ? (size_t)pl.nseg * 3 * pl.S * pl.plane_bytes : 0;
}

template <typename T>
T* at(void* ws, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(ws) + off); }

int fill(float* p, long long n, float v, hipStream_t st) {
  if (n <= 0) return 0;
  const long long blocks = std::min<long long>((n + 255) / 256, 16384);
  fill_kernel<<<(unsigned)blocks, 256, 0, st>>>(p, n, v);
  LAUNCH_CHECK();
  return 0;
}

// Validates geometry (IndexError semantics of the reference's advanced
// indexing) and uploads source cells and the receiver map.
int setup_geometry(const Plan& pl, const gbw_acoustic2d_inputs* in, void* ws, hipStream_t st,
                   std::vector<int>& host_rcell) {
  const Dims& d = pl.d;
  if (!in || !in->src_x || !in->src_z) return fail(GBW_ERR_INVALID, "source positions are NULL");
  if (pl.R > 0 && (!in->rcv_x || !in->rcv_z)) return fail(GBW_ERR_INVALID, "receiver positions are NULL");
  std::vector<int> src(pl.S);
  for (int s = 0; s < pl.S; ++s) {
    const long long z = in->src_z[s] + d.nabc, x = in->src_x[s] + d.nabc;
    if (z < 0 || z >= d.NZ || x < 0 || x >= d.NX)
      return fail(GBW_ERR_INDEX, "index out of range: source " + std::to_string(s) + " at (z=" +
                                     std::to_string(in->src_z[s]) + ", x=" +
                                     std::to_string(in->src_x[s]) + ")");
    src[s] = (int)(z * d.PX + x);
  }
  host_rcell.assign(std::max(1, pl.R), 0);
  std::vector<int> rnext(std::max(1, pl.R), -1), heads;
  std::vector<unsigned char> rrow(d.NZ, 0);
  std::unordered_map<int, int> last;  // cell -> last receiver in its list
  for (int r = 0; r < pl.R; ++r) {
    const long long z = in->rcv_z[r] + d.nabc, x = in->rcv_x[r] + d.nabc;
    // vx has NX-1 columns and vz NZ-1 rows: the reference indexes all three fields
    if (z < 0 || z >= d.NZ - 1 || x < 0 || x >= d.NX - 1)
      return fail(GBW_ERR_INDEX, "index out of range: receiver " + std::to_string(r) + " at (z=" +
                                     std::to_string(in->rcv_z[r]) + ", x=" +
                                     std::to_string(in->rcv_x[r]) + ")");
    const int cell = (int)(z * d.PX + x);
    host_rcell[r] = cell;
    rrow[z] = 1;
    auto it = last.find(cell);
    if (it == last.end()) {
      heads.push_back(cell);
      heads.push_back(r);
      last.emplace(cell, r);
    } else {
      rnext[it->second] = r;
      it->second = r;
    }
  }
  const int n_heads = (int)heads.size() / 2;
  std::vector<int> hc(std::max(1, n_heads)), hr(std::max(1, n_heads));
  for (int i = 0; i < n_heads; ++i) { hc[i] = heads[2 * i]; hr[i] = heads[2 * i + 1]; }
  int* rhead = at<int>(ws, pl.off_rhead);
  const long long blocks = std::min<long long>((d.plane + 255) / 256, 16384);
  fill_int_kernel<<<(unsigned)blocks, 256, 0, st>>>(rhead, d.plane, -1);
  LAUNCH_CHECK();
  int* heads_d = at<int>(ws, pl.off_heads);
  HIP_TRY(hipMemcpyAsync(at<int>(ws, pl.off_src), src.data(), src.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at<int>(ws, pl.off_rnext), rnext.data(), rnext.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at<int>(ws, pl.off_rcell), host_rcell.data(),
                         host_rcell.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(heads_d, hc.data(), (size_t)n_heads * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(heads_d + std::max(1, pl.R), hr.data(), (size_t)n_heads * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at<unsigned char>(ws, pl.off_rrow), rrow.data(), rrow.size(),
                         hipMemcpyHostToDevice, st));
  if (n_heads > 0) {
    rmap_scatter_kernel<<<(n_heads + 255) / 256, 256, 0, st>>>(rhead, heads_d, heads_d + std::max(1, pl.R),
                                                              n_heads);
    LAUNCH_CHECK();
  }
  // pageable copies are staged before hipMemcpyAsync returns, but keep the
  // host vectors alive until the stream has consumed them to be safe
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

int build_coefs(const Plan& pl, const gbw_acoustic2d_problem* p, const gbw_acoustic2d_inputs* in,
                void* ws, hipStream_t st) {
  if (!in->vp || !in->rho || !in->damp || !in->src_v)
    return fail(GBW_ERR_INVALID, "vp, rho, damp and src_v must be device pointers");
  const Dims& d = pl.d;
  dim3 grid((d.PX + 255) / 256, d.NZ);
  coef_kernel<<<grid, 256, 0, st>>>(d, in->vp, in->rho, in->damp, p->dt, p->dz,
                                    at<float>(ws, pl.off_coef));
  LAUNCH_CHECK();
  return 0;
}

template <int TZ, bool RECOMP>
void launch_fwd(const Plan& pl, const FwdArgs& a, hipStream_t st) {
  dim3 grid(pl.tiles_x * pl.tiles_z, pl.groups);
  fwd_step_kernel<TZ, RECOMP><<<grid, NTHREADS, 0, st>>>(a);
}

template <int TZ>
void launch_adj(const Plan& pl, const AdjArgs& a, hipStream_t st) {
  dim3 grid(pl.tiles_x * pl.tiles_z, pl.groups);
  adj_step_kernel<TZ><<<grid, NTHREADS, 0, st>>>(a);
}

FwdArgs base_fwd_args(const Plan& pl, const gbw_acoustic2d_problem* p,
                      const gbw_acoustic2d_inputs* in, void* ws) {
  FwdArgs a{};
  a.d = pl.d; a.S = pl.S; a.spb = pl.spb; a.tiles_x = pl.tiles_x;
  a.coef = at<float>(ws, pl.off_coef);
  a.src_cell = at<int>(ws, pl.off_src);
  a.wav = in->src_v;
  a.wav_stride = p->wavelet_shared ? 0 : p->nt;
  a.nt = p->nt; a.dt = p->dt;
  a.rhead = at<int>(ws, pl.off_rhead);
  a.rnext = at<int>(ws, pl.off_rnext);
  a.rrow = at<unsigned char>(ws, pl.off_rrow);
  a.R = pl.R;
  return a;
}

void set_state(FwdArgs& a, float* in, float* out, long long fstride) {
  a.P0 = in; a.VX0 = in + fstride; a.VZ0 = in + 2 * fstride;
  a.P1 = out; a.VX1 = out + fstride; a.VZ1 = out + 2 * fstride;
}

}  // namespace

extern "C" {

int gbw_abi_version(void) { return GBW_ABI_VERSION; }

const char* gbw_last_error(void) { return g_last_error.c_str(); }

size_t gbw_acoustic2d_workspace_bytes(const gbw_acoustic2d_problem* prob, int pass) {
  Plan pl;
  if (pass < 0 || pass > 2) return 0;
  if (make_plan(prob, pass, pl)) return 0;
  return pl.total;
}

size_t gbw_acoustic2d_checkpoint_bytes(const gbw_acoustic2d_problem* prob) {
  Plan pl;
  if (make_plan(prob, 0, pl)) return 0;
  return checkpoint_bytes(pl);
}

int gbw_acoustic2d_forward(const gbw_acoustic2d_problem* prob, const gbw_acoustic2d_inputs* in,
                           const gbw_acoustic2d_outputs* out, void* workspace,
                           size_t workspace_bytes, void* stream) {
  Plan pl;
  int rc = make_plan(prob, 0, pl);
  if (rc) return rc;
  if (!out) return fail(GBW_ERR_INVALID, "outputs is NULL");
  if (!workspace || workspace_bytes < pl.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(pl.total) + " bytes");
  if (pl.k && !out->checkpoints) return fail(GBW_ERR_INVALID, "checkpoint_every > 0 needs a checkpoint buffer");
  if (prob->snapshot_every > 0 && !out->snapshots) return fail(GBW_ERR_INVALID, "snapshot buffer is NULL");
  const bool rec = out->rec_p || out->rec_vx || out->rec_vz;
  if (rec && !(out->rec_p && out->rec_vx && out->rec_vz))
    return fail(GBW_ERR_INVALID, "rec_p, rec_vx and rec_vz must be given together");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<int> rcell;
  if ((rc = setup_geometry(pl, in, workspace, st, rcell))) return rc;
  if ((rc = build_coefs(pl, prob, in, workspace, st))) return rc;
  const Dims& d = pl.d;
  const long long fstride = (long long)pl.S * d.plane;
  const long long inner = (long long)d.nz * d.nx;
  const bool energy = out->energy_p || out->energy_vx || out->energy_vz;
  float* A = at<float>(workspace, pl.off_A);
  float* B = at<float>(workspace, pl.off_B);
  float* E = at<float>(workspace, pl.off_E);
  float* ck = static_cast<float*>(out->checkpoints);
  const long long slot = 3 * fstride;
  if (energy) {
    if ((rc = fill(E, (long long)pl.groups * 3 * inner, 0.f, st))) return rc;
    if (out->energy_p && (rc = fill(out->energy_p, inner, 0.f, st))) return rc;
    if (out->energy_vx && (rc = fill(out->energy_vx, inner, 0.f, st))) return rc;
    if (out->energy_vz && (rc = fill(out->energy_vz, inner, 0.f, st))) return rc;
  }
  float* cur = pl.k ? ck : A;
  if ((rc = fill(cur, slot, 0.f, st))) return rc;
  FwdArgs a = base_fwd_args(pl, prob, in, workspace);
  a.rec_p = out->rec_p; a.rec_vx = out->rec_vx; a.rec_vz = out->rec_vz;
  if (!rec) a.rec_p = a.rec_vx = a.rec_vz = nullptr;
  a.energy = energy ? E : nullptr;
  const int se = prob->snapshot_every;
  int snap = 0;
  for (int it = 0; it < pl.nt; ++it) {
    float* next;
    if (pl.k && (it + 1) % pl.k == 0 && it + 1 < pl.nt)
      next = ck + (long long)((it + 1) / pl.k) * slot;
    else
      next = (cur == A) ? B : A;
    set_state(a, cur, next, fstride);
    a.it = it;
    if (pl.TZ == 16) launch_fwd<16, false>(pl, a, st); else launch_fwd<8, false>(pl, a, st);
    LAUNCH_CHECK();
    cur = next;
    const bool chunk_end = ((it + 1) % pl.echunk == 0) || it + 1 == pl.nt;
    if (energy && chunk_end) {
      energy_flush_kernel<<<(unsigned)((3 * inner + 255) / 256), 256, 0, st>>>(
          E, pl.groups, inner, out->energy_p, out->energy_vx, out->energy_vz);
      LAUNCH_CHECK();
    }
    if (se > 0 && (((it + 1) % se == 0) || it + 1 == pl.nt)) {
      dim3 g((d.nx + 255) / 256, d.nz);
      snapshot_kernel<<<g, 256, 0, st>>>(d, cur, out->snapshots + (long long)snap * inner);
      LAUNCH_CHECK();
      ++snap;
    }
  }
  return 0;
}

int gbw_acoustic2d_backward(const gbw_acoustic2d_problem* prob, const gbw_acoustic2d_inputs* in,
                            const gbw_acoustic2d_adjoint* adj, void* workspace,
                            size_t workspace_bytes, void* stream) {
  if (!adj) return fail(GBW_ERR_INVALID, "adjoint is NULL");
  const int pass = adj->grad_rho ? 2 : 1;
  Plan pl;
  int rc = make_plan(prob, pass, pl);
  if (rc) return rc;
  if (!workspace || workspace_bytes < pl.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(pl.total) + " bytes");
  if (!adj->checkpoints) return fail(GBW_ERR_INVALID, "checkpoint buffer is NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<int> rcell;
  if ((rc = setup_geometry(pl, in, workspace, st, rcell))) return rc;
  if ((rc = build_coefs(pl, prob, in, workspace, st))) return rc;
  const Dims& d = pl.d;
  const long long fstride = (long long)pl.S * d.plane;
  const long long slot = 3 * fstride;
  float* A = at<float>(workspace, pl.off_A);
  float* B = at<float>(workspace, pl.off_B);
  float* LA = at<float>(workspace, pl.off_LA);
  float* LB = at<float>(workspace, pl.off_LB);
  float* D = at<float>(workspace, pl.off_D);
  float* GX = pass == 2 ? at<float>(workspace, pl.off_GX) : nullptr;
  float* GZ = pass == 2 ? at<float>(workspace, pl.off_GZ) : nullptr;
  float* G1 = at<float>(workspace, pl.off_G1);
  float* G2 = pass == 2 ? at<float>(workspace, pl.off_G2) : nullptr;
  float* GW = at<float>(workspace, pl.off_GW);
  const float* ck = static_cast<const float*>(adj->checkpoints);
  if ((rc = fill(G1, (long long)pl.groups * d.plane, 0.f, st))) return rc;
  if (G2 && (rc = fill(G2, (long long)pl.groups * d.plane, 0.f, st))) return rc;
  if ((rc = fill(GW, (long long)pl.S * pl.nt, 0.f, st))) return rc;
  if ((rc = fill(LA, slot, 0.f, st))) return rc;
  const bool any_g = adj->g_p || adj->g_vx || adj->g_vz;
  if (any_g && pl.R > 0) {
    adj_inject_kernel<<<(pl.S * pl.R + 255) / 256, 256, 0, st>>>(
        d, pl.S, pl.nt, pl.nt - 1, pl.R, at<int>(workspace, pl.off_rcell),
        at<int>(workspace, pl.off_rhead), at<int>(workspace, pl.off_rnext), adj->g_p, adj->g_vx,
        adj->g_vz, LA, LA + fstride, LA + 2 * fstride);
    LAUNCH_CHECK();
  }
  FwdArgs f = base_fwd_args(pl, prob, in, workspace);
  AdjArgs b{};
  b.d = d; b.S = pl.S; b.spb = pl.spb; b.tiles_x = pl.tiles_x;
  b.coef = f.coef; b.src_cell = f.src_cell; b.nt = pl.nt; b.dt = prob->dt;
  b.rhead = f.rhead; b.rnext = f.rnext; b.rrow = f.rrow; b.R = pl.R;
  b.g_p = adj->g_p; b.g_vx = adj->g_vx; b.g_vz = adj->g_vz;
  b.gA1 = G1; b.gA2 = G2; b.gw = GW;
  float* lcur = LA;
  for (int c = pl.nseg - 1; c >= 0; --c) {
    const int t0 = c * pl.k, t1 = std::min(pl.nt, t0 + pl.k);
    // recompute the interval from its checkpoint, keeping D^n (and vx/vz brackets)
    float* cur = const_cast<float*>(ck + (long long)c * slot);
    for (int it = t0; it < t1; ++it) {
      float* next = (cur == A) ? B : A;
      set_state(f, cur, next, fstride);
      f.it = it;
      f.dbuf = D + (long long)(it - t0) * fstride;
      f.gxbuf = GX ? GX + (long long)(it - t0) * fstride : nullptr;
      f.gzbuf = GZ ? GZ + (long long)(it - t0) * fstride : nullptr;
      if (pl.TZ == 16) launch_fwd<16, true>(pl, f, st); else launch_fwd<8, true>(pl, f, st);
      LAUNCH_CHECK();
      cur = next;
    }
    for (int it = t1 - 1; it >= t0; --it) {
      float* lnext = (lcur == LA) ? LB : LA;
      b.LP0 = lcur; b.LX0 = lcur + fstride; b.LZ0 = lcur + 2 * fstride;
      b.LP1 = lnext; b.LX1 = lnext + fstride; b.LZ1 = lnext + 2 * fstride;
      b.it = it;
      b.dbuf = D + (long long)(it - t0) * fstride;
      b.gxbuf = GX ? GX + (long long)(it - t0) * fstride : nullptr;
      b.gzbuf = GZ ? GZ + (long long)(it - t0) * fstride : nullptr;
      if (pl.TZ == 16) launch_adj<16>(pl, b, st); else launch_adj<8>(pl, b, st);
      LAUNCH_CHECK();
      lcur = lnext;
    }
  }
  if (adj->grad_vp || adj->grad_rho) {
    dim3 g((d.nx + 127) / 128, d.nz);
    float* gv = adj->grad_vp;
    float* tmp = nullptr;
    if (!gv) {  // fold_kernel always writes grad_vp: borrow the D buffer as a sink
      tmp = D;
      gv = tmp;
    }
    fold_kernel<<<g, 128, 0, st>>>(d, pl.groups, G1, G2, in->vp, in->rho, prob->dt, prob->dz, gv,
                                   adj->grad_rho);
    LAUNCH_CHECK();
  }
  if (adj->grad_src_v) {
    const int n = prob->wavelet_shared ? pl.nt : pl.S * pl.nt;
    wavelet_grad_kernel<<<(n + 255) / 256, 256, 0, st>>>(GW, pl.S, pl.nt, prob->wavelet_shared,
                                                         adj->grad_src_v);
    LAUNCH_CHECK();
  }
  return 0;
}

}  // extern "C"
