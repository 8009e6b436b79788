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
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "gbw_acoustic2d.h"
#include "gbw_common.h"

using namespace gbw;

namespace {

constexpr float kC1 = 9.0f / 8.0f;
constexpr float kC2 = (float)(-1.0 / 24.0);
constexpr int TX = 128;       // tile width (columns): 32 float4 groups
// Occupancy target (waves per SIMD) of the forward step kernel: 4 lets two
// 512-thread workgroups share a CU (at most 128 VGPRs).  16-row tiles fit that
// budget; 32-row tiles and the adjoint kernel spill under it and run faster
// with one workgroup per CU (measured on MI355X, 64 shots, 390 x 1740).
template <int TZ> constexpr int fwd_waves() { return TZ == 16 ? 4 : 1; }
constexpr int NTHREADS = 512; // eight 64-lane wavefronts per workgroup


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
  int S, spb, tiles_x, ntiles;   // shots, shots per workgroup, tiles along x, workgroup units
  int seg_h;                     // z-marching kernels: rows per segment
  int am_segs;                   // z-marching two-step forward: segments per column strip
  int mix_ta, mix_spb_b;         // two-phase two-step launches: phase-A tiles, phase-B group size
  const float* P0; const float* VX0; const float* VZ0;  // input state fields (shot-major)
  float* P1; float* VX1; float* VZ1;                    // output state fields
  const float* coef;
  const int* src_cell;           // [S * src_k] padded cell index (row * PX + col) of each point
  const float* wav; long long wav_stride;  // w[s * wav_stride + k * wav_kstride + it]
  int it, nt; float dt;
  const int* rhead; const int* rnext; const unsigned char* rrow; int R;
  float* rec_p; float* rec_vx; float* rec_vz;  // [S][nt][R] or null
  float* energy;                 // [group][3][nz][nx] or null
  float* dbuf;                   // [S][plane] bracket D of this step, or null
  float* gxbuf; float* gzbuf;    // [S][plane] vx / vz brackets of this step, or null
  long long bstride;             // two-step kernel: offset of step it+1's brackets
  int one;                       // two-step kernel: do step it only (the last of a problem)
  int skip1;                     // two-step kernel: step it is before the problem's first
                                 // step (it == -1): run step it+1 only, from the input state
  long long wav_kstride;         // wavelet row of point k of a shot: + k * wav_kstride
  int src_k;                     // source points per shot (distinct cells)
};

// ---------------------------------------------------------------------------
// Staging geometry shared by the step kernels
//
// A workgroup (512 threads = eight 64-lane wavefronts) owns a tile of TZ rows
// x TX = 128 columns.  Every staged row covers columns x0-4 .. x0+TX+3 as
// GW + 2 float4 groups, loaded with 16-byte global loads; in LDS a row holds
// GW + 4 float4 (columns x0-8 .. x0+TX+7) so that the 4-wide stencil reads of
// the outermost
// groups stay inside the row (guard groups 0 and 19 are zero).  Work items are
// (row, float4 group) pairs: 4 cells along x per thread and item.
// ---------------------------------------------------------------------------
constexpr int GW = TX / 4;  // float4 groups per tile row
constexpr int NG = GW + 2;  // float4 groups staged per row (one halo group each side)
constexpr int LG = GW + 4;  // float4 groups per LDS row (plus a zero guard each side)

__device__ __forceinline__ float4 f4zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ float f4get(const float4& v, int e) {
  return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}
__device__ __forceinline__ void f4set(float4& v, int e, float s) {
  if (e == 0) v.x = s; else if (e == 1) v.y = s; else if (e == 2) v.z = s; else v.w = s;
}

// float4 index of row gz, first column x (a multiple of 4), clamped into the
// plane; ok tells whether the group is real (rows outside the grid and the
// columns left of 0 read as zero; columns NX..PX-1 are zero padding in memory)
__device__ __forceinline__ unsigned group_index(const Dims& d, int gz, int x, bool& ok) {
  ok = gz >= 0 && gz < d.NZ && x >= 0 && x <= d.PX - 4;
  const int zc = min(max(gz, 0), d.NZ - 1);
  const int xc = min(max(x, 0), d.PX - 4);
  return (unsigned)((zc * d.PX + xc) >> 2);
}

// tile index of this workgroup: consecutive workgroups go to different XCDs
// (round-robin over 8), so hand each XCD a contiguous band of tiles and
// neighbouring tiles share that XCD's L2
__device__ __forceinline__ int xcd_tile() {
  const int b = blockIdx.x;
  return (b & 7) * (gridDim.x >> 3) + (b >> 3);
}

// RECOMP = false: forward pass (receivers, energy); true: checkpoint-interval
// recompute for the adjoint (stores the brackets the adjoint needs instead).
// MP: several source points per shot (src_k > 1); the one-point instance
// compiles with src_k fixed at 1
template <int TZ, bool RECOMP, bool MP = false>
__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(fwd_waves<TZ>()))) void fwd_step_kernel(FwdArgs a) {
  const int SK = MP ? a.src_k : 1;
  constexpr int RV = TZ + 3;  // vx rows z0-1 .. z0+TZ+1
  constexpr int RZ = TZ + 6;  // vz rows z0-3 .. z0+TZ+2
  constexpr int RP = TZ + 3;  // p' rows z0-1 .. z0+TZ+1
  constexpr int NV = (RV * NG + NTHREADS - 1) / NTHREADS;
  constexpr int NZZ = (RZ * NG + NTHREADS - 1) / NTHREADS;
  constexpr int NP = (RP * NG + NTHREADS - 1) / NTHREADS;
  constexpr int NT = TZ * GW / NTHREADS;  // tile items (TZ rows x GW groups)
  static_assert(TZ % 16 == 0, "TZ must be a multiple of 16");
  __shared__ float4 sVX[RV][LG];
  __shared__ float4 sVZ[RZ][LG];
  __shared__ float4 sP[RP][LG];
  const Dims d = a.d;
  const int tid = threadIdx.x;
  const int tile = xcd_tile();
  if (tile >= a.ntiles) return;
  const int x0 = (tile % a.tiles_x) * TX;
  const int z0 = (tile / a.tiles_x) * TZ;
  const int s_begin = blockIdx.y * a.spb;
  const int s_end = min(a.S, s_begin + a.spb);
  if (s_begin >= s_end) return;  // uniform per workgroup
  const float4* __restrict__ omk1 = reinterpret_cast<const float4*>(a.coef);
  const float4* __restrict__ a1 = reinterpret_cast<const float4*>(a.coef + d.plane);
  const float4* __restrict__ omk2 = reinterpret_cast<const float4*>(a.coef + 2 * d.plane);
  const float4* __restrict__ omk3 = reinterpret_cast<const float4*>(a.coef + 3 * d.plane);
  const float4* __restrict__ a2 = reinterpret_cast<const float4*>(a.coef + 4 * d.plane);

  for (int i = tid; i < RV; i += NTHREADS) sVX[i][0] = sVX[i][LG - 1] = f4zero();
  for (int i = tid; i < RZ; i += NTHREADS) sVZ[i][0] = sVZ[i][LG - 1] = f4zero();

  // --- per-workgroup constants (registers across the shot loop) ---
  unsigned vx_idx[NV], vz_idx[NZZ], p_idx[NP];
  unsigned vx_ok = 0, vz_ok = 0, p_ok = 0;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const int it = tid + n * NTHREADS, r = it / NG, g = it % NG;
    bool ok;
    vx_idx[n] = group_index(d, z0 - 1 + r, x0 - 4 + 4 * g, ok);
    if (ok && it < RV * NG) vx_ok |= 1u << n;
  }
#pragma unroll
  for (int n = 0; n < NZZ; ++n) {
    const int it = tid + n * NTHREADS, r = it / NG, g = it % NG;
    bool ok;
    vz_idx[n] = group_index(d, z0 - 3 + r, x0 - 4 + 4 * g, ok);
    if (ok && it < RZ * NG) vz_ok |= 1u << n;
  }
  // p' items: region mask per cell (4 bits per item), coefficients
  float4 pc_omk[NP], pc_a1[NP];
  unsigned p_reg = 0, p_in = 0;
#pragma unroll
  for (int n = 0; n < NP; ++n) {
    const int it = tid + n * NTHREADS, r = it / NG, g = it % NG;
    const int gz = z0 - 1 + r, xb = x0 - 4 + 4 * g;
    bool ok;
    p_idx[n] = group_index(d, gz, xb, ok);
    ok = ok && it < RP * NG;
    if (ok) p_ok |= 1u << n;
    const bool zr = gz >= d.f + 1 && gz <= d.NZ - 3;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      if (ok && x < d.NX) p_in |= 1u << (4 * n + e);
      if (ok && zr && x >= 2 && x <= d.NX - 3) p_reg |= 1u << (4 * n + e);
    }
    pc_omk[n] = omk1[p_idx[n]];
    pc_a1[n] = a1[p_idx[n]];
  }
  // tile items
  unsigned t_idx[NT];
  float4 tc_omk2[NT], tc_omk3[NT], tc_a2[NT];
  unsigned t_ok = 0, t_x = 0, t_z = 0, t_e = 0, t_rcv = 0;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const int it = tid + n * NTHREADS, tr = it / GW, h = it % GW;
    const int gz = z0 + tr, xb = x0 + 4 * h;
    bool ok;
    t_idx[n] = group_index(d, gz, xb, ok);
    if (ok) t_ok |= 1u << n;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      const bool in = ok && x < d.NX;
      if (in && gz >= d.f && gz <= d.NZ - 2 && x >= 1 && x <= d.NX - 3) t_x |= 1u << (4 * n + e);
      if (in && gz >= d.f && gz <= d.NZ - 3 && x >= 1 && x <= d.NX - 2) t_z |= 1u << (4 * n + e);
      if (in && gz >= d.nabc && gz < d.nabc + d.nz && x >= d.nabc && x < d.nabc + d.nx)
        t_e |= 1u << (4 * n + e);
    }
    // receivers on this row, or (free surface, row f) vz samples of row f-1,
    // which the mirror vz[f-1] = vz[f] fills from this row
    if (!RECOMP && a.rec_p && ok && (a.rrow[gz] || (d.fs && gz == d.f && a.rrow[gz - 1])))
      t_rcv |= 1u << n;
    tc_omk2[n] = omk2[t_idx[n]];
    tc_omk3[n] = omk3[t_idx[n]];
    tc_a2[n] = a2[t_idx[n]];
  }

  float4 e_p[NT], e_vx[NT], e_vz[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) e_p[n] = e_vx[n] = e_vz[n] = f4zero();

  float4 rvx[NV], rvz[NZZ], rp[NP];
  auto load_v = [&](int s) {
    const long long so = (long long)s * d.plane;
    const float4* __restrict__ VX = reinterpret_cast<const float4*>(a.VX0 + so);
    const float4* __restrict__ VZ = reinterpret_cast<const float4*>(a.VZ0 + so);
#pragma unroll
    for (int n = 0; n < NV; ++n) rvx[n] = VX[vx_idx[n]];
#pragma unroll
    for (int n = 0; n < NZZ; ++n) rvz[n] = VZ[vz_idx[n]];
  };
  auto load_p = [&](int s) {
    const float4* __restrict__ P = reinterpret_cast<const float4*>(a.P0 + (long long)s * d.plane);
#pragma unroll
    for (int n = 0; n < NP; ++n) rp[n] = P[p_idx[n]];
  };
  load_v(s_begin);
  load_p(s_begin);

  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const int scell = a.src_cell[s * SK];
    const float supd = a.dt * a.wav[(long long)s * a.wav_stride + a.it];
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const int it = tid + n * NTHREADS;
      if (it < RV * NG) sVX[it / NG][it % NG + 1] = ((vx_ok >> n) & 1u) ? rvx[n] : f4zero();
    }
#pragma unroll
    for (int n = 0; n < NZZ; ++n) {
      const int it = tid + n * NTHREADS;
      if (it < RZ * NG) sVZ[it / NG][it % NG + 1] = ((vz_ok >> n) & 1u) ? rvz[n] : f4zero();
    }
    if (s + 1 < s_end) load_v(s + 1);  // next shot's vx/vz loads overlap this shot
    __syncthreads();
    // p' on the tile and its halo (acoustic_2d.py:213-237)
#pragma unroll
    for (int n = 0; n < NP; ++n) {
      const int it = tid + n * NTHREADS;
      if (it < RP * NG) {
        const int r = it / NG, g = it % NG;
        const float4 xa = sVX[r][g], xb = sVX[r][g + 1], xc = sVX[r][g + 2];
        const float vxs[12] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w};
        const float4 zm2 = sVZ[r][g + 1], zm1 = sVZ[r + 1][g + 1];
        const float4 z00 = sVZ[r + 2][g + 1], zp1 = sVZ[r + 3][g + 1];
        const float4 p0 = ((p_ok >> n) & 1u) ? rp[n] : f4zero();
        float4 out = p0, dv = f4zero();
        const unsigned base = p_idx[n] * 4u;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const unsigned bit = 1u << (4 * n + e);
          float pv = f4get(p0, e);
          float br = 0.f;
          if (p_reg & bit) {
            float s1 = vxs[4 + e] - vxs[3 + e];
            s1 = s1 + f4get(z00, e);
            s1 = s1 - f4get(zm1, e);
            float s2 = vxs[5 + e] - vxs[2 + e];
            s2 = s2 + f4get(zp1, e);
            s2 = s2 - f4get(zm2, e);
            const float t1 = kC1 * s1;
            const float t2 = kC2 * s2;
            br = t1 + t2;
            const float u = f4get(pc_omk[n], e) * pv;
            pv = u - f4get(pc_a1[n], e) * br;
          }
          if ((p_in & bit) && (int)(base + e) == scell) pv = pv + supd;
          // further points of the shot (acoustic_2d.py:234-237; distinct
          // cells, so at most one adds here)
          if (MP && (p_in & bit))
            for (int k = 1; k < SK; ++k)
              if ((int)(base + e) == a.src_cell[s * SK + k])
                pv = pv + a.dt * a.wav[(long long)s * a.wav_stride + k * a.wav_kstride + a.it];
          f4set(out, e, pv);
          f4set(dv, e, br);
        }
        sP[r][g + 1] = out;
        if constexpr (RECOMP) {
          // tile groups only: rows 1..TZ of the rectangle, float4 groups 1..GW
          if (r >= 1 && r <= TZ && g >= 1 && g <= GW && ((p_ok >> n) & 1u))
            reinterpret_cast<float4*>(a.dbuf + so)[p_idx[n]] = dv;
        }
      }
    }
    if (s + 1 < s_end) load_p(s + 1);
    __syncthreads();
    if (d.fs) {
      const int rm = d.f - z0;  // rect row of f-1; f+1 is rect row rm+2
      if (rm >= 0 && rm <= TZ) {
        for (int g = tid; g < NG; g += NTHREADS) {
          const float4 v = sP[rm + 2][g + 1];
          sP[rm][g + 1] = make_float4(-v.x, -v.y, -v.z, -v.w);
        }
      }
      __syncthreads();
    }
    // v' on the tile (acoustic_2d.py:243-274)
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      if ((t_ok >> n) & 1u) {
        const int it = tid + n * NTHREADS, tr = it / GW, h = it % GW;
        const int gz = z0 + tr;
        const float4 pa = sP[tr + 1][h + 1], pb = sP[tr + 1][h + 2], pcc = sP[tr + 1][h + 3];
        const float ps[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w, pcc.x, pcc.y, pcc.z, pcc.w};
        const float4 pu = sP[tr][h + 2], pd1 = sP[tr + 2][h + 2], pd2 = sP[tr + 3][h + 2];
        float4 vx = sVX[tr + 1][h + 2];
        float4 vz = sVZ[tr + 3][h + 2];
        float4 gxv = f4zero(), gzv = f4zero();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const unsigned bit = 1u << (4 * n + e);
          if (t_x & bit) {
            const float d1 = ps[5 + e] - ps[4 + e];
            const float d2 = ps[6 + e] - ps[3 + e];
            const float br = kC1 * d1 + kC2 * d2;
            const float u = f4get(tc_omk2[n], e) * f4get(vx, e);
            f4set(vx, e, u - f4get(tc_a2[n], e) * br);
            f4set(gxv, e, br);
          }
          if (t_z & bit) {
            const float d1 = f4get(pd1, e) - f4get(pb, e);
            const float d2 = f4get(pd2, e) - f4get(pu, e);
            const float br = kC1 * d1 + kC2 * d2;
            const float u = f4get(tc_omk3[n], e) * f4get(vz, e);
            f4set(vz, e, u - f4get(tc_a2[n], e) * br);
            f4set(gzv, e, br);
          }
        }
        const unsigned ti = t_idx[n];
        reinterpret_cast<float4*>(a.P1 + so)[ti] = pb;
        reinterpret_cast<float4*>(a.VX1 + so)[ti] = vx;
        if (!(d.fs && gz == d.f - 1)) reinterpret_cast<float4*>(a.VZ1 + so)[ti] = vz;
        if (d.fs && gz == d.f) reinterpret_cast<float4*>(a.VZ1 + so)[ti - (d.PX >> 2)] = vz;
        if constexpr (RECOMP) {
          if (a.gxbuf) {
            reinterpret_cast<float4*>(a.gxbuf + so)[ti] = gxv;
            reinterpret_cast<float4*>(a.gzbuf + so)[ti] = gzv;
          }
        } else {
          const bool rcv = (t_rcv >> n) & 1u;
          const int4 rh = rcv ? reinterpret_cast<const int4*>(a.rhead)[ti] : make_int4(-1, -1, -1, -1);
          const bool mirror_row = d.fs && gz == d.f - 1;  // its vz sample comes from row f
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            for (int hh = e == 0 ? rh.x : (e == 1 ? rh.y : (e == 2 ? rh.z : rh.w)); hh >= 0;
                 hh = a.rnext[hh]) {
              const long long o = ((long long)s * a.nt + a.it) * a.R + hh;
              a.rec_p[o] = f4get(pb, e);
              a.rec_vx[o] = f4get(vx, e);
              if (!mirror_row) a.rec_vz[o] = f4get(vz, e);
            }
          }
          if (rcv && d.fs && gz == d.f) {
            const int4 ru = reinterpret_cast<const int4*>(a.rhead)[ti - (d.PX >> 2)];
#pragma unroll
            for (int e = 0; e < 4; ++e)
              for (int hh = e == 0 ? ru.x : (e == 1 ? ru.y : (e == 2 ? ru.z : ru.w)); hh >= 0;
                   hh = a.rnext[hh])
                a.rec_vz[((long long)s * a.nt + a.it) * a.R + hh] = f4get(vz, e);
          }
          e_p[n] = make_float4(e_p[n].x + pb.x * pb.x, e_p[n].y + pb.y * pb.y,
                               e_p[n].z + pb.z * pb.z, e_p[n].w + pb.w * pb.w);
          e_vx[n] = make_float4(e_vx[n].x + vx.x * vx.x, e_vx[n].y + vx.y * vx.y,
                                e_vx[n].z + vx.z * vx.z, e_vx[n].w + vx.w * vx.w);
          e_vz[n] = make_float4(e_vz[n].x + vz.x * vz.x, e_vz[n].y + vz.y * vz.y,
                                e_vz[n].z + vz.z * vz.z, e_vz[n].w + vz.w * vz.w);
        }
      }
    }
    __syncthreads();
  }
  if constexpr (!RECOMP) {
    if (a.energy) {
      const long long inner = (long long)d.nz * d.nx;
      float* eg = a.energy + (long long)blockIdx.y * 3 * inner;
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int it = tid + n * NTHREADS, tr = it / GW, h = it % GW;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if ((t_e >> (4 * n + e)) & 1u) {
            const long long o = (long long)(z0 + tr - d.nabc) * d.nx + (x0 + 4 * h + e - d.nabc);
            eg[o] = eg[o] + f4get(e_p[n], e);
            eg[inner + o] = eg[inner + o] + f4get(e_vx[n], e);
            eg[2 * inner + o] = eg[2 * inner + o] + f4get(e_vz[n], e);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Two-step forward kernel (temporal blocking)
//
// One launch advances every shot by steps it and it+1, so the state is read
// and written once per two steps.  A workgroup stages its TZ x TX tile with a
// 6-cell halo (rows z0-6 .. z0+TZ+5; float4 groups -2 .. GW+1, i.e. columns
// x0-8 .. x0+TX+7) once per shot, then computes in LDS, relative to the tile
// (rows, float4 groups; 0 = first tile row / group):
//   p'   rows -4 .. TZ+4, groups -1 .. GW+1   (reads vx groups -2 .. GW+2, vz rows -6 .. TZ+5)
//   v'   rows -3 .. TZ+2, groups -1 .. GW
//   p''  rows -1 .. TZ+1, groups -1 .. GW
//   v''  the tile
// which are the halos step it+1 of the tile needs (the stencil reaches 2 cells
// back and 1 forward from p to v, 1 back and 2 forward from v to p).  Cells
// outside those regions hold stale values that nothing reads.  Each cell
// belongs to one thread ("slot") in every phase, so fields are updated in
// place; the free-surface mirrors run between phases.  Per cell, the
// arithmetic is the one-step kernel's, in the same order.
// ---------------------------------------------------------------------------
constexpr int W2 = GW + 4;   // staged float4 groups per row (-2 .. GW+1)
constexpr int MAX_SPB = 64;  // shots per workgroup of the two-step kernel
#ifndef FWD2_THREADS
#define FWD2_THREADS 1024     // one staged group per thread for 16-row tiles
#endif
constexpr int LW2 = GW + 6;  // LDS row: staged groups plus a zero guard each side

// per-slot flag word: bits 0-6 slot flags, then 4-bit per-cell masks
constexpr unsigned F_OK = 1u, F_PV = 2u, F_P1 = 4u, F_V1 = 8u, F_P2 = 16u, F_T = 32u, F_RCV = 64u;
constexpr unsigned F_PFULL = 128u, F_VFULL = 1u << 28;  // all 4 cells in the P / X and Z regions
constexpr unsigned F_AINT = 1u << 29;  // adjoint: no region, grid or free-surface exception
constexpr unsigned F_RMORE = 1u << 30;  // forward: a cell of the slot has several receivers
// two-phase two-step launches: shots per phase-B group (groups of 4, 8, 16
// and 32 measured 339, 315, 323 and 332 us per forward launch on 64 shots of
// 390 x 1740 with phase-A groups of 32 or 64)
constexpr int MIX_SPB_B = 8;
constexpr int M_PREG = 8, M_XREG = 12, M_ZREG = 16, M_PIN = 20, M_ENRG = 24;

// one workgroup's tile `tile` and shots [s_begin, s_end) of a problem; the
// forward's energy sums go to partial map `epart` of a.energy
// PHASED: the batched gradient's launches, in which a problem may start at the
// second step (a.skip1); the other kernels compile without that test
template <int TZ, int TXT, int NTH, bool RECOMP, bool PHASED = false>
__device__ __forceinline__ void fwd2_tile(const FwdArgs& a, int tile, int s_begin, int s_end,
                                          int epart) {
  // this kernel's tile width (columns) and the staging geometry that follows
  constexpr int TX = TXT, GW = TXT / 4, W2 = GW + 4, LW2 = GW + 6;
  constexpr int SR = TZ + 12;                    // staged rows (-6 .. TZ+5)
  constexpr int NS = (SR * W2 + NTH - 1) / NTH;  // slots per thread
  // LDS row index = tile row + 6; LDS column of staged group g (0 = group -2) is g + 1
  __shared__ float4 sP[SR][LW2];
  __shared__ float4 sVX[SR][LW2];
  __shared__ float4 sVZ[SR][LW2];
  float4 eacc[RECOMP ? 1 : 3][NS];  // energy sums of the slot's cells (tile slots)
  const Dims d = a.d;
  const int tid = threadIdx.x;
  const int x0 = (tile % a.tiles_x) * TX;
  const int z0 = (tile / a.tiles_x) * TZ;
  if (s_begin >= s_end) return;  // uniform per workgroup
  const float4* __restrict__ omk1 = reinterpret_cast<const float4*>(a.coef);
  const float4* __restrict__ a1 = reinterpret_cast<const float4*>(a.coef + d.plane);
  const float4* __restrict__ omk2 = reinterpret_cast<const float4*>(a.coef + 2 * d.plane);
  const float4* __restrict__ omk3 = reinterpret_cast<const float4*>(a.coef + 3 * d.plane);
  const float4* __restrict__ a2 = reinterpret_cast<const float4*>(a.coef + 4 * d.plane);

  for (int i = tid; i < SR; i += NTH) {
    sP[i][0] = sP[i][LW2 - 1] = f4zero();
    sVX[i][0] = sVX[i][LW2 - 1] = f4zero();
    sVZ[i][0] = sVZ[i][LW2 - 1] = f4zero();
  }

  // --- per-slot constants (registers across the shot loop) ---
  unsigned idx[NS], fl[NS];
  float4 c_omk1[NS], c_a1[NS], c_omk2[NS], c_omk3[NS], c_a2[NS];
  // forward: first receiver of each cell of a receiver slot (-1: none), by slot
  __shared__ int4 sRH[RECOMP ? 1 : SR * W2];
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const int item = tid + n * NTH;
    const int r = item / W2, g = item % W2;
    const int gz = z0 - 6 + r, xb = x0 - 8 + 4 * g;
    bool ok;
    idx[n] = group_index(d, gz, xb, ok);
    const bool valid = item < SR * W2;
    ok = ok && valid;
    const bool p1 = valid && r >= 2 && r <= TZ + 10 && g >= 1;
    const bool v1 = valid && r >= 3 && r <= TZ + 8 && g >= 1 && g <= GW + 2;
    const bool p2 = valid && r >= 5 && r <= TZ + 7 && g >= 1 && g <= GW + 2;
    const bool tt = valid && r >= 6 && r <= TZ + 5 && g >= 2 && g <= GW + 1;
    unsigned f = 0;
    if (ok) f |= F_OK;
    if (ok && r >= 2 && r <= TZ + 10) f |= F_PV;
    if (p1) f |= F_P1;
    if (v1) f |= F_V1;
    if (p2) f |= F_P2;
    if (tt) f |= F_T;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      const bool in = ok && x < d.NX;
      if (in) f |= 1u << (M_PIN + e);
      if (in && gz >= d.f + 1 && gz <= d.NZ - 3 && x >= 2 && x <= d.NX - 3) f |= 1u << (M_PREG + e);
      if (in && gz >= d.f && gz <= d.NZ - 2 && x >= 1 && x <= d.NX - 3) f |= 1u << (M_XREG + e);
      if (in && gz >= d.f && gz <= d.NZ - 3 && x >= 1 && x <= d.NX - 2) f |= 1u << (M_ZREG + e);
      if (tt && in && gz >= d.nabc && gz < d.nabc + d.nz && x >= d.nabc && x < d.nabc + d.nx)
        f |= 1u << (M_ENRG + e);
    }
    if (!RECOMP && a.rec_p && tt && ok && (a.rrow[gz] || (d.fs && gz == d.f && a.rrow[gz - 1])))
      f |= F_RCV;
    if (((f >> M_PREG) & 15u) == 15u) f |= F_PFULL;
    if (((f >> M_XREG) & 15u) == 15u && ((f >> M_ZREG) & 15u) == 15u) f |= F_VFULL;
    if constexpr (!RECOMP) {
      // the slot's receivers: the first of each cell to LDS, and whether a
      // cell has more
      if (f & F_RCV) {
        const int4 rh = reinterpret_cast<const int4*>(a.rhead)[idx[n]];
        sRH[item] = rh;
        if ((rh.x >= 0 && a.rnext[rh.x] >= 0) || (rh.y >= 0 && a.rnext[rh.y] >= 0) ||
            (rh.z >= 0 && a.rnext[rh.z] >= 0) || (rh.w >= 0 && a.rnext[rh.w] >= 0))
          f |= F_RMORE;
      }
    }
    fl[n] = f;
    c_omk1[n] = p1 ? omk1[idx[n]] : f4zero();
    c_a1[n] = p1 ? a1[idx[n]] : f4zero();
    c_omk2[n] = v1 ? omk2[idx[n]] : f4zero();
    c_omk3[n] = v1 ? omk3[idx[n]] : f4zero();
    c_a2[n] = v1 ? a2[idx[n]] : f4zero();
    if constexpr (!RECOMP) {
      eacc[0][n] = eacc[1][n] = eacc[2][n] = f4zero();
    }
  }

  // p and vx are needed on the p' rows only; the other slots load the tile's
  // first group (one shared line per wavefront) instead of branching around
  // the load, which keeps the three prefetch loads back to back
  unsigned pv_idx[NS];
  const unsigned idx0 = (unsigned)((min(z0, d.NZ - 1) * d.PX + x0) >> 2);
#pragma unroll
  for (int n = 0; n < NS; ++n) pv_idx[n] = (fl[n] & F_PV) ? idx[n] : idx0;
  float4 rp[NS], rvx[NS], rvz[NS];
  auto load = [&](int s) {
    const long long so = (long long)s * d.plane;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(a.P0 + so);
    const float4* __restrict__ VX = reinterpret_cast<const float4*>(a.VX0 + so);
    const float4* __restrict__ VZ = reinterpret_cast<const float4*>(a.VZ0 + so);
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      rp[n] = P[pv_idx[n]];
      rvx[n] = VX[pv_idx[n]];
      rvz[n] = VZ[idx[n]];
    }
  };
  load(s_begin);

  // p-update of the slot's group (acoustic_2d.py:213-237); returns the new
  // group, the bracket D in dv
  // each slot's own p, vx, vz live in registers across the phases; LDS holds
  // them for the neighbours' stencils
  float4 own_p[NS], own_vx[NS], own_vz[NS];
  // p-update of the slot's group (acoustic_2d.py:213-237): computed for all
  // four cells, then cells outside the P-region keep p (a select, not a
  // branch; whole groups inside the region skip the selects).  The source
  // term is added when the shot's source lies in the staged region.
  auto p_update = [&](int n, int r, int c, bool src_here, int scell, float supd,
                      float4& dv) -> float4 {
    const float4 xa = sVX[r][c - 1], xb = own_vx[n], xc = sVX[r][c + 1];
    const float vxs[12] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w};
    const float4 zm2 = sVZ[r - 2][c], zm1 = sVZ[r - 1][c];
    const float4 z00 = own_vz[n], zp1 = sVZ[r + 1][c];
    const float4 p0 = own_p[n];
    float4 out, brv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float s1 = vxs[4 + e] - vxs[3 + e];
      s1 = s1 + f4get(z00, e);
      s1 = s1 - f4get(zm1, e);
      float s2 = vxs[5 + e] - vxs[2 + e];
      s2 = s2 + f4get(zp1, e);
      s2 = s2 - f4get(zm2, e);
      const float t1 = kC1 * s1;
      const float t2 = kC2 * s2;
      const float br = t1 + t2;
      const float u = f4get(c_omk1[n], e) * f4get(p0, e);
      f4set(out, e, u - f4get(c_a1[n], e) * br);
      f4set(brv, e, br);
    }
    if (fl[n] & F_PFULL) {
      dv = brv;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool m = (fl[n] >> (M_PREG + e)) & 1u;
        f4set(out, e, m ? f4get(out, e) : f4get(p0, e));
        f4set(dv, e, m ? f4get(brv, e) : 0.f);
      }
    }
    if (src_here) {
      const unsigned base = idx[n] * 4u;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (((fl[n] >> (M_PIN + e)) & 1u) && (int)(base + e) == scell)
          f4set(out, e, f4get(out, e) + supd);
    }
    return out;
  };
  // v-update of the slot's group (acoustic_2d.py:243-265) from the p in sP;
  // cells outside the X / Z regions keep their value
  auto v_update = [&](int n, int r, int c, float4& vx, float4& vz, float4& gxv, float4& gzv) {
    const float4 pa = sP[r][c - 1], pb = own_p[n], pcc = sP[r][c + 1];
    const float ps[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w, pcc.x, pcc.y, pcc.z, pcc.w};
    const float4 pu = sP[r - 1][c], pd1 = sP[r + 1][c], pd2 = sP[r + 2][c];
    const float4 vx0 = own_vx[n], vz0 = own_vz[n];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      {
        const float d1 = ps[5 + e] - ps[4 + e];
        const float d2 = ps[6 + e] - ps[3 + e];
        const float br = kC1 * d1 + kC2 * d2;
        const float u = f4get(c_omk2[n], e) * f4get(vx0, e);
        f4set(vx, e, u - f4get(c_a2[n], e) * br);
        f4set(gxv, e, br);
      }
      {
        const float d1 = f4get(pd1, e) - f4get(pb, e);
        const float d2 = f4get(pd2, e) - f4get(pu, e);
        const float br = kC1 * d1 + kC2 * d2;
        const float u = f4get(c_omk3[n], e) * f4get(vz0, e);
        f4set(vz, e, u - f4get(c_a2[n], e) * br);
        f4set(gzv, e, br);
      }
    }
    if (!(fl[n] & F_VFULL)) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool mx = (fl[n] >> (M_XREG + e)) & 1u;
        const bool mz = (fl[n] >> (M_ZREG + e)) & 1u;
        f4set(vx, e, mx ? f4get(vx, e) : f4get(vx0, e));
        f4set(gxv, e, mx ? f4get(gxv, e) : 0.f);
        f4set(vz, e, mz ? f4get(vz, e) : f4get(vz0, e));
        f4set(gzv, e, mz ? f4get(gzv, e) : 0.f);
      }
    }
  };
  // the energy maps are sums of squares over shots and steps: one fused
  // multiply-add per term (they agree with the reference's separately rounded
  // products to rounding, as the two-step association already does)
  auto add_sq = [](float4& acc, const float4& v) {
    acc = make_float4(__builtin_fmaf(v.x, v.x, acc.x), __builtin_fmaf(v.y, v.y, acc.y),
                      __builtin_fmaf(v.z, v.z, acc.z), __builtin_fmaf(v.w, v.w, acc.w));
  };
  // receivers of a tile slot: p, vx, vz of step `step`; for the free-surface
  // row f the slot also writes the vz samples of row f-1 (the mirror's source).
  // The first receiver of each cell comes from registers, so a shot's samples
  // are stores only (no dependent loads in front of the next barrier); cells
  // with more receivers walk their list.
  auto record = [&](int n, int s, int step, int gz, const float4& pv, const float4& vx, const float4& vz) {
    const unsigned ti = idx[n];
    const int4 rh = sRH[RECOMP ? 0 : tid + n * NTH];
    const bool mirror_row = d.fs && gz == d.f - 1;
    const long long ob = ((long long)s * a.nt + step) * a.R;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int hh = e == 0 ? rh.x : (e == 1 ? rh.y : (e == 2 ? rh.z : rh.w));
      if (hh < 0) continue;
      a.rec_p[ob + hh] = f4get(pv, e);
      a.rec_vx[ob + hh] = f4get(vx, e);
      if (!mirror_row) a.rec_vz[ob + hh] = f4get(vz, e);
      if (fl[n] & F_RMORE) {
        for (hh = a.rnext[hh]; hh >= 0; hh = a.rnext[hh]) {
          a.rec_p[ob + hh] = f4get(pv, e);
          a.rec_vx[ob + hh] = f4get(vx, e);
          if (!mirror_row) a.rec_vz[ob + hh] = f4get(vz, e);
        }
      }
    }
    if (d.fs && gz == d.f) {
      const int4 ru = reinterpret_cast<const int4*>(a.rhead)[ti - (d.PX >> 2)];
#pragma unroll
      for (int e = 0; e < 4; ++e)
        for (int hh = e == 0 ? ru.x : (e == 1 ? ru.y : (e == 2 ? ru.z : ru.w)); hh >= 0;
             hh = a.rnext[hh])
          a.rec_vz[ob + hh] = f4get(vz, e);
    }
  };

  const int rf = d.f - 1 - (z0 - 6);  // LDS row of the free-surface mirror row f-1
  // per-shot source cell and source terms of this workgroup's shots, staged in
  // LDS up front: reading them in the shot loop then never waits on the
  // in-flight state prefetch
  __shared__ int sSrc[MAX_SPB];
  __shared__ float sW[2][MAX_SPB];
  for (int i = tid; i < s_end - s_begin; i += NTH) {
    const float* wv = a.wav + (long long)(s_begin + i) * a.wav_stride + a.it;
    sSrc[i] = a.src_cell[s_begin + i];
    sW[0][i] = (PHASED && a.skip1) ? 0.f : a.dt * wv[0];  // step it exists only without `skip1`
    sW[1][i] = a.one ? 0.f : a.dt * wv[1];    // step it+1 exists only without `one`
  }
  __syncthreads();
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const int scell = sSrc[s - s_begin];
    const float supd0 = sW[0][s - s_begin];
    const float supd1 = sW[1][s - s_begin];
    const int sz = scell / d.PX, sx = scell - sz * d.PX;
    const bool src_here = sz >= z0 - 6 && sz < z0 + TZ + 6 && sx >= x0 - 8 && sx < x0 + TX + 8;
    // stage this shot
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const int item = tid + n * NTH;
      if (item < SR * W2) {
        const int r = item / W2, c = item % W2 + 1;
        const bool ok = fl[n] & F_OK;
        own_p[n] = (fl[n] & F_PV) ? rp[n] : f4zero();
        own_vx[n] = (fl[n] & F_PV) ? rvx[n] : f4zero();
        own_vz[n] = ok ? rvz[n] : f4zero();
        sP[r][c] = own_p[n];
        sVX[r][c] = own_vx[n];
        sVZ[r][c] = own_vz[n];
      }
    }
    if (s + 1 < s_end) load(s + 1);  // next shot's loads overlap this shot
    __syncthreads();
    // step it (uniform per launch: a problem that starts at step it+1 skips it;
    // its staged input state then feeds step it+1 directly)
    if (!(PHASED && a.skip1)) {
    // step it: p'
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      if (fl[n] & F_P1) {
        const int item = tid + n * NTH;
        const int r = item / W2, c = item % W2 + 1;
        float4 dv;
        const float4 pn = p_update(n, r, c, src_here, scell, supd0, dv);
        own_p[n] = pn;
        sP[r][c] = pn;  // nothing reads sP in this phase
        if constexpr (RECOMP) {
          if ((fl[n] & (F_T | F_OK)) == (F_T | F_OK))
            reinterpret_cast<float4*>(a.dbuf + so)[idx[n]] = dv;
        }
      }
    }
    __syncthreads();
    if (d.fs && rf >= 2 && rf + 2 <= TZ + 10) {  // p'[f-1] = -p'[f+1]
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const int item = tid + n * NTH;
        if (item < SR * W2 && item / W2 == rf) {
          const float4 v = sP[rf + 2][item % W2 + 1];
          own_p[n] = make_float4(-v.x, -v.y, -v.z, -v.w);
          sP[rf][item % W2 + 1] = own_p[n];
        }
      }
      __syncthreads();
    }
    // step it: v'
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      if (fl[n] & F_V1) {
        const int item = tid + n * NTH;
        const int r = item / W2, c = item % W2 + 1;
        float4 vx, vz, gxv, gzv;
        v_update(n, r, c, vx, vz, gxv, gzv);
        own_vx[n] = vx;
        own_vz[n] = vz;
        // nothing reads sVX / sVZ in this phase
        sVX[r][c] = vx;
        sVZ[r][c] = vz;
        if constexpr (RECOMP) {
          if (a.gxbuf && (fl[n] & (F_T | F_OK)) == (F_T | F_OK)) {
            reinterpret_cast<float4*>(a.gxbuf + so)[idx[n]] = gxv;
            reinterpret_cast<float4*>(a.gzbuf + so)[idx[n]] = gzv;
          }
        }
      }
    }
    __syncthreads();
    if (d.fs && rf >= 3 && rf + 1 <= TZ + 8) {  // vz'[f-1] = vz'[f]
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const int item = tid + n * NTH;
        if (item < SR * W2 && item / W2 == rf) {
          own_vz[n] = sVZ[rf + 1][item % W2 + 1];
          sVZ[rf][item % W2 + 1] = own_vz[n];
        }
      }
      __syncthreads();
    }
    // step it: receivers and energy of the tile
    if constexpr (!RECOMP) {
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        if ((fl[n] & (F_T | F_OK)) == (F_T | F_OK)) {
          const int item = tid + n * NTH;
          const int r = item / W2;
          const float4 pv = own_p[n], vx = own_vx[n], vz = own_vz[n];
          if (fl[n] & F_RCV) record(n, s, a.it, z0 - 6 + r, pv, vx, vz);
          if (a.energy) {
            add_sq(eacc[0][n], pv);
            add_sq(eacc[1][n], vx);
            add_sq(eacc[2][n], vz);
          }
        }
      }
    }
    }  // !skip1
    if (a.one) {
      // step it only: store p', v' of the tile (LDS holds both mirrors'
      // results for every tile row, row f-1 included)
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        if ((fl[n] & (F_T | F_OK)) == (F_T | F_OK)) {
          const unsigned ti = idx[n];
          reinterpret_cast<float4*>(a.P1 + so)[ti] = own_p[n];
          reinterpret_cast<float4*>(a.VX1 + so)[ti] = own_vx[n];
          reinterpret_cast<float4*>(a.VZ1 + so)[ti] = own_vz[n];
        }
      }
      __syncthreads();
      continue;
    }
    // step it+1: p'' (reads only v' and this slot's own p')
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      if (fl[n] & F_P2) {
        const int item = tid + n * NTH;
        const int r = item / W2, c = item % W2 + 1;
        float4 dv;
        const float4 pn = p_update(n, r, c, src_here, scell, supd1, dv);
        own_p[n] = pn;
        sP[r][c] = pn;
        if constexpr (RECOMP) {
          if ((fl[n] & (F_T | F_OK)) == (F_T | F_OK))
            reinterpret_cast<float4*>(a.dbuf + a.bstride + so)[idx[n]] = dv;
        }
      }
    }
    __syncthreads();
    if (d.fs && rf >= 5 && rf + 2 <= TZ + 7) {  // p''[f-1] = -p''[f+1]
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const int item = tid + n * NTH;
        if (item < SR * W2 && item / W2 == rf) {
          const float4 v = sP[rf + 2][item % W2 + 1];
          own_p[n] = make_float4(-v.x, -v.y, -v.z, -v.w);
          sP[rf][item % W2 + 1] = own_p[n];
        }
      }
      __syncthreads();
    }
    // step it+1: v'' on the tile, store the state
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      if ((fl[n] & (F_T | F_OK)) == (F_T | F_OK)) {
        const int item = tid + n * NTH;
        const int r = item / W2, c = item % W2 + 1;
        const int gz = z0 - 6 + r;
        float4 vx, vz, gxv, gzv;
        v_update(n, r, c, vx, vz, gxv, gzv);
        const float4 pv = own_p[n];
        const unsigned ti = idx[n];
        reinterpret_cast<float4*>(a.P1 + so)[ti] = pv;
        reinterpret_cast<float4*>(a.VX1 + so)[ti] = vx;
        if (!(d.fs && gz == d.f - 1)) reinterpret_cast<float4*>(a.VZ1 + so)[ti] = vz;
        if (d.fs && gz == d.f) reinterpret_cast<float4*>(a.VZ1 + so)[ti - (d.PX >> 2)] = vz;
        if constexpr (RECOMP) {
          if (a.gxbuf) {
            reinterpret_cast<float4*>(a.gxbuf + a.bstride + so)[ti] = gxv;
            reinterpret_cast<float4*>(a.gzbuf + a.bstride + so)[ti] = gzv;
          }
        } else {
          if (fl[n] & F_RCV) record(n, s, a.it + 1, gz, pv, vx, vz);
          if (a.energy) {
            add_sq(eacc[0][n], pv);
            add_sq(eacc[1][n], vx);
            add_sq(eacc[2][n], vz);
          }
        }
      }
    }
    __syncthreads();
  }
  if constexpr (!RECOMP) {
    if (a.energy) {
      const long long inner = (long long)d.nz * d.nx;
      float* eg = a.energy + (long long)epart * 3 * inner;
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        if (fl[n] & F_T) {
          const int item = tid + n * NTH;
          const int r = item / W2, g = item % W2;
          const int gz = z0 - 6 + r, xb = x0 - 8 + 4 * g;
          // nabc and nx multiples of 4 (xb is): a group wholly in the interior
          // is one aligned float4 of each map
          if (((d.nabc | d.nx) & 3) == 0 && ((fl[n] >> M_ENRG) & 15u) == 15u) {
            float4* e4 = reinterpret_cast<float4*>(eg + (long long)(gz - d.nabc) * d.nx + (xb - d.nabc));
            const long long i4 = inner >> 2;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
              const float4 o = e4[m * i4], v = eacc[m][n];
              e4[m * i4] = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
            }
            continue;
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if ((fl[n] >> (M_ENRG + e)) & 1u) {
              const long long o = (long long)(gz - d.nabc) * d.nx + (xb + e - d.nabc);
              eg[o] = eg[o] + f4get(eacc[0][n], e);
              eg[inner + o] = eg[inner + o] + f4get(eacc[1][n], e);
              eg[2 * inner + o] = eg[2 * inner + o] + f4get(eacc[2][n], e);
            }
          }
        }
      }
    }
  }
}

// shots of shot group `group` (a.spb per group)
template <int TZ, int TXT, int NTH, bool RECOMP, bool PHASED = false>
__device__ __forceinline__ void fwd2_group(const FwdArgs& a, int tile, int group) {
  fwd2_tile<TZ, TXT, NTH, RECOMP, PHASED>(a, tile, group * a.spb, min(a.S, (group + 1) * a.spb),
                                          group);
}

template <int TZ, int TXT, int NTH, bool RECOMP>
__global__ __launch_bounds__(NTH) void fwd2_step_kernel(FwdArgs a) {
  const int tile = xcd_tile();
  if (tile >= a.ntiles) return;
  fwd2_group<TZ, TXT, NTH, RECOMP>(a, tile, blockIdx.y);
}

// Two-step forward / recompute in two phases of work units, so the launch
// ends in full rounds of one workgroup per CU without giving up large shot
// groups.  Phase A: mix_ta tiles (a multiple of the CU count) in groups of
// a.spb shots, which take whole rounds; phase B: the other tiles in groups of
// mix_spb_b shots, small units that fill the last rounds evenly.  Phase A
// takes the first tiles, except with a free surface (a.d.fs): then the first
// tile rows hold the tiles that run the mirrors' extra barriers, and as phase B
// they are spread over every XCD in small units instead of lengthening one
// XCD's share of phase A (forward 339 -> 317 us per launch, 64 shots of
// 390 x 1740; without a free surface this order measured 317 against 310).
// Workgroups go to the 8 XCDs round-robin (blockIdx.x & 7); each XCD runs an
// equal share of both phases, phase A first, and within a phase its units in
// (group, tile) order over a contiguous range of tiles, so the workgroups
// resident on an XCD step neighbouring tiles through the same shots at about
// the same time and share the tiles' halos in that XCD's L2.  Energy: unit
// (group g, tile t) adds to partial map g (each map cell has one writer).
inline int mix_spb_b(int spb) { return std::min(spb, MIX_SPB_B); }

template <int TZ, int TXT, int NTH, bool RECOMP>
__global__ __launch_bounds__(NTH) void fwd2_mixed_kernel(FwdArgs a) {
  const int k = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int ta8 = a.mix_ta >> 3;                   // phase-A tiles per XCD
  const int ga = (a.S + a.spb - 1) / a.spb;        // phase-A groups
  const int na8 = ta8 * ga;                        // phase-A units per XCD
  int tile, s0, s1, g;
  const int tb = a.ntiles - a.mix_ta;              // phase-B tiles
  const int a0 = a.d.fs ? tb : 0, b0 = a.d.fs ? 0 : a.mix_ta;  // first tile of each phase
  if (j < na8) {
    g = j / ta8;
    tile = a0 + k * ta8 + j % ta8;
    s0 = g * a.spb;
    s1 = min(a.S, s0 + a.spb);
  } else {
    const int gb = (a.S + a.mix_spb_b - 1) / a.mix_spb_b;
    const int nb8 = (tb * gb + 7) / 8;             // phase-B units per XCD
    const int u = k * nb8 + (j - na8);
    if (u >= tb * gb) return;  // uniform per workgroup
    g = u / tb;
    tile = b0 + u % tb;
    s0 = g * a.mix_spb_b;
    s1 = min(a.S, s0 + a.mix_spb_b);
  }
  fwd2_tile<TZ, TXT, NTH, RECOMP>(a, tile, s0, s1, g);
}

// ---------------------------------------------------------------------------
// Batched forward: many independent problems (uneven grids, step counts and
// shot counts) advanced by one launch per two time steps.  Problems are sorted
// by step count, longest first, so the problems still running at step `it`
// are a prefix; work units (tile, shot group) of those problems are numbered
// consecutively (unit_start) and handed out XCD-aware like single-problem tiles.
// ---------------------------------------------------------------------------
struct BatchDesc {
  FwdArgs a;               // geometry, coefficients, receivers, outputs; it / state set per launch
  float* A; float* B;      // the problem's two state buffers ([field][shot][plane])
  long long fstride;       // shots * plane
  int groups;              // shot groups (units = ntiles * groups)
  int pad;
};

__global__ void batch_start_kernel(unsigned long long* done) {
  if (threadIdx.x == 0 && blockIdx.x == 0) done[0] = wall_clock64();
}

// zero the initial state of every problem (blockIdx.y = problem)
__global__ void batch_zero_kernel(const BatchDesc* __restrict__ descs) {
  const BatchDesc& b = descs[blockIdx.y];
  const long long n = 3 * b.fstride;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b.A[i] = 0.f;
}

template <int TZ, int TXT>
__global__ __launch_bounds__(FWD2_THREADS) void fwd2_batch_kernel(
    const BatchDesc* __restrict__ descs, const int* __restrict__ unit_start, int n_active, int it,
    int parity, int single, unsigned* counters, unsigned long long* done) {
  const int u = xcd_tile();
  if (u >= unit_start[n_active]) return;
  int lo = 0, hi = n_active - 1;  // last problem whose first unit is <= u
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (unit_start[mid] <= u) lo = mid; else hi = mid - 1;
  }
  const int p = lo;
  const BatchDesc& bd = descs[p];
  FwdArgs a = bd.a;
  float* in = parity ? bd.B : bd.A;
  float* out = parity ? bd.A : bd.B;
  a.P0 = in; a.VX0 = in + bd.fstride; a.VZ0 = in + 2 * bd.fstride;
  a.P1 = out; a.VX1 = out + bd.fstride; a.VZ1 = out + 2 * bd.fstride;
  a.it = it;
  a.one = single || a.nt - it == 1;
  const int local = u - unit_start[p];
  fwd2_group<TZ, TXT, FWD2_THREADS, false>(a, local % a.ntiles, local / a.ntiles);
  if (done && it + (a.one ? 1 : 2) >= a.nt) {  // the problem's last launch
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      if (atomicAdd(&counters[p], 1u) == (unsigned)(a.ntiles * bd.groups) - 1u)
        done[1 + p] = wall_clock64();
    }
  }
}

// ---------------------------------------------------------------------------
// z-marching recompute step (checkpoint-interval recompute for the adjoint)
//
// A workgroup owns one 64-column strip, one z segment [zs, ze) and one shot
// and walks down the segment TB rows at a time.  Staged rows live in LDS ring
// buffers (RING rows, indexed by row mod RING), so a row loaded for one block
// of rows is reused by the next instead of being re-read as a halo: every
// field row is read once per segment (plus the 2 float4 x-halo groups), not
// once per tile as in fwd_step_kernel.  p' runs two rows ahead of v' (v'(i)
// needs p' up to row i+2).  Workgroups are ordered (segment, strip, shot) with
// the shot fastest and a contiguous range per XCD, so the shots of a strip
// run together and share the coefficient rows in that XCD's L2.
// ---------------------------------------------------------------------------
constexpr int MB = 16;    // rows per march iteration
constexpr int RING = 32;  // LDS ring rows (live span is at most MB + 3)

__global__ __launch_bounds__(NTHREADS) void recomp_march_kernel(FwdArgs a) {
  __shared__ float4 rVX[RING][LG];
  __shared__ float4 rVZ[RING][LG];
  __shared__ float4 rP[RING][LG];
  const Dims d = a.d;
  const int tid = threadIdx.x;
  const int logical = xcd_tile();
  if (logical >= a.ntiles) return;
  const int s = logical % a.S;
  const int ts = logical / a.S;
  const int x0 = (ts % a.tiles_x) * TX;
  const int zs = (ts / a.tiles_x) * a.seg_h;
  const int ze = min(d.NZ, zs + a.seg_h);
  const long long so = (long long)s * d.plane;
  const float4* __restrict__ VX0 = reinterpret_cast<const float4*>(a.VX0 + so);
  const float4* __restrict__ VZ0 = reinterpret_cast<const float4*>(a.VZ0 + so);
  const float4* __restrict__ P0 = reinterpret_cast<const float4*>(a.P0 + so);
  const float4* __restrict__ omk1 = reinterpret_cast<const float4*>(a.coef);
  const float4* __restrict__ a1 = reinterpret_cast<const float4*>(a.coef + d.plane);
  const float4* __restrict__ omk2 = reinterpret_cast<const float4*>(a.coef + 2 * d.plane);
  const float4* __restrict__ omk3 = reinterpret_cast<const float4*>(a.coef + 3 * d.plane);
  const float4* __restrict__ a2 = reinterpret_cast<const float4*>(a.coef + 4 * d.plane);
  const int scell = a.src_cell[s];
  const float supd = a.dt * a.wav[(long long)s * a.wav_stride + a.it];
  for (int i = tid; i < RING; i += NTHREADS) rVX[i][0] = rVX[i][LG - 1] = f4zero();
  auto slot = [&](int row) { return (row - zs + RING) & (RING - 1); };

  // staged rows of one field: items (row, group) = (r0 + it / NG, it % NG)
  struct Rows {
    float4 v[2];
    unsigned idx[2];
    unsigned ok;
  };
  auto issue = [&](const float4* __restrict__ src, int r0, int n, Rows& R) {
    R.ok = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int it = tid + k * NTHREADS;
      bool ok;
      R.idx[k] = group_index(d, r0 + it / NG, x0 - 4 + 4 * (it % NG), ok);
      if (ok && it < n * NG) R.ok |= 1u << k;
      R.v[k] = src[R.idx[k]];
    }
  };
  auto commit = [&](float4 (*ring)[LG], int r0, int n, const Rows& R) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int it = tid + k * NTHREADS;
      if (it < n * NG) ring[slot(r0 + it / NG)][it % NG + 1] = ((R.ok >> k) & 1u) ? R.v[k] : f4zero();
    }
  };
  // coefficients of p' items
  struct PCoef {
    float4 omk[2], a1[2];
  };
  auto issue_pcoef = [&](const Rows& R, PCoef& C) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      C.omk[k] = omk1[R.idx[k]];
      C.a1[k] = a1[R.idx[k]];
    }
  };
  // p' for rows [r0, r0+n) from the vx/vz rings and the staged p rows
  auto p_pass = [&](int r0, int n, const Rows& Pr, const PCoef& C) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int it = tid + k * NTHREADS;
      if (it < n * NG) {
        const int row = r0 + it / NG, g = it % NG;
        const int xb = x0 - 4 + 4 * g;
        const bool ok = (Pr.ok >> k) & 1u;
        const int sr = slot(row);
        const float4 xa = rVX[sr][g], xb4 = rVX[sr][g + 1], xc = rVX[sr][g + 2];
        const float vxs[12] = {xa.x, xa.y, xa.z, xa.w, xb4.x, xb4.y, xb4.z, xb4.w,
                               xc.x, xc.y, xc.z, xc.w};
        const float4 zm2 = rVZ[slot(row - 2)][g + 1], zm1 = rVZ[slot(row - 1)][g + 1];
        const float4 z00 = rVZ[sr][g + 1], zp1 = rVZ[slot(row + 1)][g + 1];
        const float4 p0 = ok ? Pr.v[k] : f4zero();
        const bool zr = row >= d.f + 1 && row <= d.NZ - 3;
        float4 out = p0, dv = f4zero();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int x = xb + e;
          float pv = f4get(p0, e);
          if (ok && zr && x >= 2 && x <= d.NX - 3) {
            float s1 = vxs[4 + e] - vxs[3 + e];
            s1 = s1 + f4get(z00, e);
            s1 = s1 - f4get(zm1, e);
            float s2 = vxs[5 + e] - vxs[2 + e];
            s2 = s2 + f4get(zp1, e);
            s2 = s2 - f4get(zm2, e);
            const float t1 = kC1 * s1;
            const float t2 = kC2 * s2;
            const float br = t1 + t2;
            const float u = f4get(C.omk[k], e) * pv;
            pv = u - f4get(C.a1[k], e) * br;
            f4set(dv, e, br);
          }
          if (ok && x < d.NX && (int)(Pr.idx[k] * 4u + e) == scell) pv = pv + supd;
          f4set(out, e, pv);
        }
        rP[sr][g + 1] = out;
        if (ok && row >= zs && row < ze && g >= 1 && g <= GW)
          reinterpret_cast<float4*>(a.dbuf + so)[Pr.idx[k]] = dv;
      }
    }
  };
  // free surface: p'[f-1] = -p'[f+1] once row f+1 is computed and row f-1 is live
  auto mirror = [&](int r0, int n) {
    const int rf = d.f + 1;
    if (d.fs && rf >= r0 && rf < r0 + n && d.f - 1 >= zs - 1) {
      __syncthreads();
      for (int g = tid; g < NG; g += NTHREADS) {
        const float4 v = rP[slot(rf)][g + 1];
        rP[slot(d.f - 1)][g + 1] = make_float4(-v.x, -v.y, -v.z, -v.w);
      }
    }
  };
  // v' tile item: row v0 + tid / GW, group h = tid % GW (columns x0 + 4h ..)
  const int tr = tid / GW, h = tid % GW;
  struct VCoef {
    float4 omk2, omk3, a2;
    unsigned idx;
    bool ok;
  };
  auto issue_vcoef = [&](int v0, VCoef& C) {
    bool ok;
    C.idx = group_index(d, v0 + tr, x0 + 4 * h, ok);
    C.ok = ok && v0 + tr < ze;
    C.omk2 = omk2[C.idx];
    C.omk3 = omk3[C.idx];
    C.a2 = a2[C.idx];
  };
  auto v_pass = [&](int v0, const VCoef& C) {
    if (!C.ok) return;
    const int gz = v0 + tr;
    const float4 pa = rP[slot(gz)][h + 1], pb = rP[slot(gz)][h + 2], pcc = rP[slot(gz)][h + 3];
    const float ps[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w, pcc.x, pcc.y, pcc.z, pcc.w};
    const float4 pu = rP[slot(gz - 1)][h + 2], pd1 = rP[slot(gz + 1)][h + 2];
    const float4 pd2 = rP[slot(gz + 2)][h + 2];
    float4 vx = rVX[slot(gz)][h + 2];
    float4 vz = rVZ[slot(gz)][h + 2];
    float4 gxv = f4zero(), gzv = f4zero();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = x0 + 4 * h + e;
      const bool in = x < d.NX;
      if (in && gz >= d.f && gz <= d.NZ - 2 && x >= 1 && x <= d.NX - 3) {
        const float d1 = ps[5 + e] - ps[4 + e];
        const float d2 = ps[6 + e] - ps[3 + e];
        const float br = kC1 * d1 + kC2 * d2;
        const float u = f4get(C.omk2, e) * f4get(vx, e);
        f4set(vx, e, u - f4get(C.a2, e) * br);
        f4set(gxv, e, br);
      }
      if (in && gz >= d.f && gz <= d.NZ - 3 && x >= 1 && x <= d.NX - 2) {
        const float d1 = f4get(pd1, e) - f4get(pb, e);
        const float d2 = f4get(pd2, e) - f4get(pu, e);
        const float br = kC1 * d1 + kC2 * d2;
        const float u = f4get(C.omk3, e) * f4get(vz, e);
        f4set(vz, e, u - f4get(C.a2, e) * br);
        f4set(gzv, e, br);
      }
    }
    reinterpret_cast<float4*>(a.P1 + so)[C.idx] = pb;
    reinterpret_cast<float4*>(a.VX1 + so)[C.idx] = vx;
    if (!(d.fs && gz == d.f - 1)) reinterpret_cast<float4*>(a.VZ1 + so)[C.idx] = vz;
    if (d.fs && gz == d.f) reinterpret_cast<float4*>(a.VZ1 + so)[C.idx - (d.PX >> 2)] = vz;
    if (a.gxbuf) {
      reinterpret_cast<float4*>(a.gxbuf + so)[C.idx] = gxv;
      reinterpret_cast<float4*>(a.gzbuf + so)[C.idx] = gzv;
    }
  };

  // prologue: p' rows zs-1 .. zs+1 (vx, p rows zs-1..zs+1; vz rows zs-3..zs+2)
  Rows Rx, Rz, Rp;
  PCoef Cp;
  issue(VX0, zs - 1, 3, Rx);
  issue(VZ0, zs - 3, 6, Rz);
  issue(P0, zs - 1, 3, Rp);
  issue_pcoef(Rp, Cp);
  commit(rVX, zs - 1, 3, Rx);
  commit(rVZ, zs - 3, 6, Rz);
  __syncthreads();
  p_pass(zs - 1, 3, Rp, Cp);
  mirror(zs - 1, 3);
  // first iteration's rows
  issue(VX0, zs + 2, MB, Rx);
  issue(VZ0, zs + 3, MB, Rz);
  issue(P0, zs + 2, MB, Rp);
  issue_pcoef(Rp, Cp);
  VCoef Cv;
  issue_vcoef(zs, Cv);
  __syncthreads();
  for (int v0 = zs; v0 < ze; v0 += MB) {
    const int pr0 = v0 + 2;  // p' rows of this iteration: [v0+2, v0+2+MB)
    commit(rVX, pr0, MB, Rx);
    commit(rVZ, v0 + 3, MB, Rz);
    __syncthreads();
    p_pass(pr0, MB, Rp, Cp);
    VCoef Cvn = Cv;
    if (v0 + MB < ze) {  // prefetch the next iteration's rows
      issue(VX0, pr0 + MB, MB, Rx);
      issue(VZ0, v0 + 3 + MB, MB, Rz);
      issue(P0, pr0 + MB, MB, Rp);
      issue_pcoef(Rp, Cp);
      issue_vcoef(v0 + MB, Cvn);
    }
    mirror(pr0, MB);
    __syncthreads();
    v_pass(v0, Cv);
    Cv = Cvn;
    __syncthreads();
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
  int S, spb, tiles_x, ntiles;
  const float* LP0; const float* LX0; const float* LZ0;  // adjoint of state it+1 (records of it injected)
  float* LP1; float* LX1; float* LZ1;                    // adjoint of state it (records of it-1 injected)
  const float* coef;
  const float* dbuf;                       // [S][plane] bracket D^it
  const float* gxbuf; const float* gzbuf;  // [S][plane] vx/vz brackets of step it (rho gradient) or null
  const float* dbuf1;                      // two-step kernel: D^(it-1)
  const float* gxbuf1; const float* gzbuf1;  // two-step kernel: brackets of step it-1 or null
  const int* src_cell;
  int it, nt; float dt;
  const int* rhead; const int* rnext; const unsigned char* rrow; int R;
  const float* g_p; const float* g_vx; const float* g_vz;  // [S][nt][R] or null
  float* gA1;   // [group][plane]
  float* gA2;   // [group][plane] or null
  float* gw;    // [S][nt] or null
  int one;      // two-step kernel: adjoint step it only (it == 0: the problem's first step)
  int am_segs, am_seg_h;  // z-marching adjoint: z segments per strip, rows per segment
  int am_seg0_h;          // rows of its first segment (the others: am_seg_h)
  // the batched gradient's launches (PHASED kernel): the per-problem arguments
  // come from these descriptors, the rest of this struct is unused
  const struct GradDesc* batch_descs;
  const int* batch_units;  // n + 1 prefix sums of work units
  int batch_n, batch_g, batch_parity, batch_c, batch_k;
  int src_k;  // source points per shot (src_cell[s * src_k + k]; gw row s * src_k + k)
};

// RHO: also accumulate the density gradient (needs the recompute's vx/vz brackets)
// MP: several source points per shot (src_k > 1); the one-point instance
// compiles with src_k fixed at 1
template <int TZ, bool RHO, bool MP = false>
__global__ __launch_bounds__(NTHREADS) void adj_step_kernel(AdjArgs a) {
  const int SK = MP ? a.src_k : 1;
  constexpr int RV = TZ + 3;  // lambda_vx rows z0-1 ..
  constexpr int RZ = TZ + 6;  // lambda_vz rows z0-3 ..
  constexpr int RP = TZ + 3;  // lambda_p' rows z0-1 ..
  constexpr int NV = (RV * NG + NTHREADS - 1) / NTHREADS;
  constexpr int NZZ = (RZ * NG + NTHREADS - 1) / NTHREADS;
  constexpr int NP = (RP * NG + NTHREADS - 1) / NTHREADS;
  constexpr int NT = TZ * GW / NTHREADS;
  static_assert(TZ % 16 == 0, "TZ must be a multiple of 16");
  __shared__ float4 sLX[RV][LG];  // lambda_vx as read
  __shared__ float4 sQX[RV][LG];  // a2 * lambda_vx on the vx region, else 0
  __shared__ float4 sLZ[RZ][LG];  // lambda_vz, the vz[f-1] = vz[f] transpose folded in
  __shared__ float4 sQZ[RZ][LG];  // a2 * lambda_vz on the vz region, else 0
  __shared__ float4 sLP[RP][LG];  // lambda_p'
  __shared__ float4 sR[RP][LG];   // a1 * lambda_p' on the p region, else 0
  const Dims d = a.d;
  const int tid = threadIdx.x;
  const int tile = xcd_tile();
  if (tile >= a.ntiles) return;
  const int x0 = (tile % a.tiles_x) * TX;
  const int z0 = (tile / a.tiles_x) * TZ;
  const int s_begin = blockIdx.y * a.spb;
  const int s_end = min(a.S, s_begin + a.spb);
  if (s_begin >= s_end) return;
  const float4* __restrict__ omk1 = reinterpret_cast<const float4*>(a.coef);
  const float4* __restrict__ a1 = reinterpret_cast<const float4*>(a.coef + d.plane);
  const float4* __restrict__ omk2 = reinterpret_cast<const float4*>(a.coef + 2 * d.plane);
  const float4* __restrict__ omk3 = reinterpret_cast<const float4*>(a.coef + 3 * d.plane);
  const float4* __restrict__ a2 = reinterpret_cast<const float4*>(a.coef + 4 * d.plane);

  for (int i = tid; i < RV; i += NTHREADS) sQX[i][0] = sQX[i][LG - 1] = f4zero();

  // --- per-workgroup constants ---
  unsigned x_idx[NV], z_idx[NZZ], p_idx[NP];
  unsigned x_ok = 0, z_ok = 0, p_ok = 0, x_reg = 0, z_reg = 0, p_reg = 0;
  float4 xa2[NV], za2[NZZ], pa1[NP];
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const int it = tid + n * NTHREADS, r = it / NG, g = it % NG;
    const int gz = z0 - 1 + r, xb = x0 - 4 + 4 * g;
    bool ok;
    x_idx[n] = group_index(d, gz, xb, ok);
    ok = ok && it < RV * NG;
    if (ok) x_ok |= 1u << n;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      if (ok && gz >= d.f && gz <= d.NZ - 2 && x >= 1 && x <= d.NX - 3) x_reg |= 1u << (4 * n + e);
    }
    xa2[n] = a2[x_idx[n]];
  }
#pragma unroll
  for (int n = 0; n < NZZ; ++n) {
    const int it = tid + n * NTHREADS, r = it / NG, g = it % NG;
    const int gz = z0 - 3 + r, xb = x0 - 4 + 4 * g;
    bool ok;
    z_idx[n] = group_index(d, gz, xb, ok);
    ok = ok && it < RZ * NG;
    if (ok) z_ok |= 1u << n;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      if (ok && gz >= d.f && gz <= d.NZ - 3 && x >= 1 && x <= d.NX - 2) z_reg |= 1u << (4 * n + e);
    }
    za2[n] = a2[z_idx[n]];
  }
#pragma unroll
  for (int n = 0; n < NP; ++n) {
    const int it = tid + n * NTHREADS, r = it / NG, g = it % NG;
    const int gz = z0 - 1 + r, xb = x0 - 4 + 4 * g;
    bool ok;
    p_idx[n] = group_index(d, gz, xb, ok);
    ok = ok && it < RP * NG;
    if (ok) p_ok |= 1u << n;
    const bool zr = gz >= d.f + 1 && gz <= d.NZ - 3;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      if (ok && zr && x >= 2 && x <= d.NX - 3) p_reg |= 1u << (4 * n + e);
    }
    pa1[n] = a1[p_idx[n]];
  }
  unsigned t_idx[NT];
  float4 tc_omk1[NT], tc_omk2[NT], tc_omk3[NT];
  unsigned t_ok = 0, t_p = 0, t_x = 0, t_z = 0, t_vx = 0, t_rcv = 0;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const int it = tid + n * NTHREADS, tr = it / GW, h = it % GW;
    const int gz = z0 + tr, xb = x0 + 4 * h;
    bool ok;
    t_idx[n] = group_index(d, gz, xb, ok);
    if (ok) t_ok |= 1u << n;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      const bool in = ok && x < d.NX;
      const unsigned bit = 1u << (4 * n + e);
      if (in && gz >= d.f + 1 && gz <= d.NZ - 3 && x >= 2 && x <= d.NX - 3) t_p |= bit;
      if (in && gz >= d.f && gz <= d.NZ - 2 && x >= 1 && x <= d.NX - 3) t_x |= bit;
      if (in && gz >= d.f && gz <= d.NZ - 3 && x >= 1 && x <= d.NX - 2) t_z |= bit;
      if (in && x <= d.NX - 2) t_vx |= bit;  // vx has NX-1 columns
    }
    if (ok && a.it >= 1 && a.rrow[gz]) t_rcv |= 1u << n;
    tc_omk1[n] = omk1[t_idx[n]];
    tc_omk2[n] = omk2[t_idx[n]];
    tc_omk3[n] = omk3[t_idx[n]];
  }
  // the free-surface fold touches this tile's vz rectangle at rect row f - (z0-3)
  const int rfold = d.f - (z0 - 3);
  const bool fold = d.fs && rfold >= 0 && rfold < RZ;

  float4 gacc[NT], gacc2[RHO ? NT : 1];
#pragma unroll
  for (int n = 0; n < NT; ++n) gacc[n] = f4zero();
  if constexpr (RHO) {
#pragma unroll
    for (int n = 0; n < NT; ++n) gacc2[n] = f4zero();
  }

  float4 rlx[NV], rlz[NZZ], rlp[NP], rd[NT];
  auto load_v = [&](int s) {
    const long long so = (long long)s * d.plane;
    const float4* __restrict__ LX = reinterpret_cast<const float4*>(a.LX0 + so);
    const float4* __restrict__ LZ = reinterpret_cast<const float4*>(a.LZ0 + so);
#pragma unroll
    for (int n = 0; n < NV; ++n) rlx[n] = LX[x_idx[n]];
#pragma unroll
    for (int n = 0; n < NZZ; ++n) rlz[n] = LZ[z_idx[n]];
  };
  auto load_p = [&](int s) {
    const long long so = (long long)s * d.plane;
    const float4* __restrict__ LP = reinterpret_cast<const float4*>(a.LP0 + so);
    const float4* __restrict__ DB = reinterpret_cast<const float4*>(a.dbuf + so);
#pragma unroll
    for (int n = 0; n < NP; ++n) rlp[n] = LP[p_idx[n]];
#pragma unroll
    for (int n = 0; n < NT; ++n) rd[n] = DB[t_idx[n]];
  };
  load_v(s_begin);
  load_p(s_begin);

  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const int it = tid + n * NTHREADS;
      if (it < RV * NG) {
        const float4 v = ((x_ok >> n) & 1u) ? rlx[n] : f4zero();
        float4 q = f4zero();
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if ((x_reg >> (4 * n + e)) & 1u) f4set(q, e, f4get(xa2[n], e) * f4get(v, e));
        sLX[it / NG][it % NG + 1] = v;
        sQX[it / NG][it % NG + 1] = q;
      }
    }
#pragma unroll
    for (int n = 0; n < NZZ; ++n) {
      const int it = tid + n * NTHREADS;
      if (it < RZ * NG) {
        const float4 v = ((z_ok >> n) & 1u) ? rlz[n] : f4zero();
        sLZ[it / NG][it % NG + 1] = v;
        if (!fold) {
          float4 q = f4zero();
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if ((z_reg >> (4 * n + e)) & 1u) f4set(q, e, f4get(za2[n], e) * f4get(v, e));
          sQZ[it / NG][it % NG + 1] = q;
        }
      }
    }
    if (s + 1 < s_end) load_v(s + 1);
    __syncthreads();
    if (fold) {  // transpose of vz[f-1] = vz[f]: lambda_vz[f] += lambda_vz[f-1]; lambda_vz[f-1] = 0
      for (int g = tid; g < NG; g += NTHREADS) {
        float4 up;
        if (rfold >= 1) {
          up = sLZ[rfold - 1][g + 1];
          sLZ[rfold - 1][g + 1] = f4zero();
        } else {
          bool ok;
          const unsigned gi = group_index(d, d.f - 1, x0 - 4 + 4 * g, ok);
          up = ok ? reinterpret_cast<const float4*>(a.LZ0 + so)[gi] : f4zero();
        }
        const float4 v = sLZ[rfold][g + 1];
        sLZ[rfold][g + 1] = make_float4(v.x + up.x, v.y + up.y, v.z + up.z, v.w + up.w);
      }
      __syncthreads();
#pragma unroll
      for (int n = 0; n < NZZ; ++n) {
        const int it = tid + n * NTHREADS;
        if (it < RZ * NG) {
          const float4 v = sLZ[it / NG][it % NG + 1];
          float4 q = f4zero();
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if ((z_reg >> (4 * n + e)) & 1u) f4set(q, e, f4get(za2[n], e) * f4get(v, e));
          sQZ[it / NG][it % NG + 1] = q;
        }
      }
      __syncthreads();
    }
    // lambda_p' on the tile and its halo
#pragma unroll
    for (int n = 0; n < NP; ++n) {
      const int it = tid + n * NTHREADS;
      if (it < RP * NG) {
        const int r = it / NG, g = it % NG;
        const int gz = z0 - 1 + r;
        const float4 qa = sQX[r][g], qb = sQX[r][g + 1], qc = sQX[r][g + 2];
        const float qx[12] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w, qc.x, qc.y, qc.z, qc.w};
        const float4 zm2 = sQZ[r][g + 1], zm1 = sQZ[r + 1][g + 1];
        const float4 z00 = sQZ[r + 2][g + 1], zp1 = sQZ[r + 3][g + 1];
        const bool ok = (p_ok >> n) & 1u;
        const float4 lp0 = ok ? rlp[n] : f4zero();
        float4 lpm4 = f4zero();
        if (d.fs && ok && gz == d.f + 1)  // lambda_p at row f-1, same columns
          lpm4 = reinterpret_cast<const float4*>(a.LP0 + so)[p_idx[n] - (d.PX >> 2) * 2];
        float4 lpv = f4zero(), rrv = f4zero();
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float lp = 0.f, rr = 0.f;
          if (ok) {
            lp = f4get(lp0, e);
            const float az = kC1 * (f4get(z00, e) - f4get(zm1, e)) + kC2 * (f4get(zp1, e) - f4get(zm2, e));
            const float ax = kC1 * (qx[4 + e] - qx[3 + e]) + kC2 * (qx[5 + e] - qx[2 + e]);
            float acc = az;
            acc = acc + ax;
            lp = lp + acc;
            if (d.fs) {
              if (gz == d.f + 1) {
                // lambda_p'[f-1] = lambda_p[f-1] + c2 * q_z[f]: rows above f carry no q
                const float zero = 0.f;
                const float azm = kC1 * (zero - zero) + kC2 * (f4get(zm1, e) - zero);
                const float axm = kC1 * (zero - zero) + kC2 * (zero - zero);
                float accm = azm;
                accm = accm + axm;
                const float lpm = f4get(lpm4, e) + accm;
                lp = lp - lpm;
              } else if (gz == d.f - 1) {
                lp = 0.f;
              }
            }
            if ((p_reg >> (4 * n + e)) & 1u) rr = f4get(pa1[n], e) * lp;
          }
          f4set(lpv, e, lp);
          f4set(rrv, e, rr);
        }
        sLP[r][g + 1] = lpv;
        sR[r][g + 1] = rrv;
      }
    }
    float4 dcur[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) dcur[n] = rd[n];
    if (s + 1 < s_end) load_p(s + 1);
    __syncthreads();
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      if ((t_ok >> n) & 1u) {
        const int it = tid + n * NTHREADS, tr = it / GW, h = it % GW;
        const int gz = z0 + tr;
        const unsigned ti = t_idx[n];
        const float4 lp4 = sLP[tr + 1][h + 2];
        const float4 ra = sR[tr + 1][h + 1], rb = sR[tr + 1][h + 2], rc = sR[tr + 1][h + 3];
        const float rs[12] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w, rc.x, rc.y, rc.z, rc.w};
        const float4 ru = sR[tr][h + 2], rd1 = sR[tr + 2][h + 2], rd2 = sR[tr + 3][h + 2];
        float4 lx4 = sLX[tr + 1][h + 2];
        float4 lz4 = sLZ[tr + 3][h + 2];
        float4 gxv = f4zero(), gzv = f4zero();
        if constexpr (RHO) {
          gxv = reinterpret_cast<const float4*>(a.gxbuf + so)[ti];
          gzv = reinterpret_cast<const float4*>(a.gzbuf + so)[ti];
        }
        const int4 rh = ((t_rcv >> n) & 1u) ? reinterpret_cast<const int4*>(a.rhead)[ti]
                                           : make_int4(-1, -1, -1, -1);
        float4 lpo = lp4;
        const int scell = a.src_cell[s * SK];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const unsigned bit = 1u << (4 * n + e);
          const float lp = f4get(lp4, e);
          if (t_p & bit) {
            f4set(gacc[n], e, f4get(gacc[n], e) - lp * f4get(dcur[n], e));
            f4set(lpo, e, f4get(tc_omk1[n], e) * lp);
          }
          if ((int)(ti * 4u + e) == scell && a.gw && (t_ok >> n & 1u) && x0 + 4 * h + e < d.NX)
            a.gw[(long long)s * SK * a.nt + a.it] = a.dt * lp;
          if (MP && a.gw && (t_ok >> n & 1u) && x0 + 4 * h + e < d.NX)
            for (int k = 1; k < SK; ++k)
              if ((int)(ti * 4u + e) == a.src_cell[s * SK + k])
                a.gw[((long long)s * SK + k) * a.nt + a.it] = a.dt * lp;
          float lx = f4get(lx4, e), lz = f4get(lz4, e);
          if constexpr (RHO) {
            float g2 = f4get(gacc2[n], e);
            if (t_z & bit) g2 = g2 - lz * f4get(gzv, e);
            if (t_x & bit) g2 = g2 - lx * f4get(gxv, e);
            f4set(gacc2[n], e, g2);
          }
          if (t_x & bit) lx = f4get(tc_omk2[n], e) * lx;
          if (t_z & bit) lz = f4get(tc_omk3[n], e) * lz;
          if (t_vx & bit)
            lx += kC1 * (rs[5 + e] - rs[4 + e]) + kC2 * (rs[6 + e] - rs[3 + e]);
          if (gz <= d.NZ - 2)
            lz += kC1 * (f4get(rd1, e) - f4get(rb, e)) + kC2 * (f4get(rd2, e) - f4get(ru, e));
          float lpv = f4get(lpo, e);
          for (int hh = e == 0 ? rh.x : (e == 1 ? rh.y : (e == 2 ? rh.z : rh.w)); hh >= 0;
               hh = a.rnext[hh]) {
            const long long o = ((long long)s * a.nt + (a.it - 1)) * a.R + hh;
            if (a.g_p) lpv += a.g_p[o];
            if (a.g_vx) lx += a.g_vx[o];
            if (a.g_vz) lz += a.g_vz[o];
          }
          f4set(lpo, e, lpv);
          f4set(lx4, e, lx);
          f4set(lz4, e, lz);
        }
        reinterpret_cast<float4*>(a.LP1 + so)[ti] = lpo;
        reinterpret_cast<float4*>(a.LX1 + so)[ti] = lx4;
        reinterpret_cast<float4*>(a.LZ1 + so)[ti] = lz4;
      }
    }
    __syncthreads();
  }
  float4* g1 = reinterpret_cast<float4*>(a.gA1 + (long long)blockIdx.y * d.plane);
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    if ((t_ok >> n) & 1u) {
      const float4 v = g1[t_idx[n]];
      g1[t_idx[n]] = make_float4(v.x + gacc[n].x, v.y + gacc[n].y, v.z + gacc[n].z, v.w + gacc[n].w);
      if constexpr (RHO) {
        float4* g2 = reinterpret_cast<float4*>(a.gA2 + (long long)blockIdx.y * d.plane);
        const float4 w = g2[t_idx[n]];
        g2[t_idx[n]] = make_float4(w.x + gacc2[n].x, w.y + gacc2[n].y, w.z + gacc2[n].z,
                                   w.w + gacc2[n].w);
      }
    }
  }
}

// Descriptor of one problem of the batched gradient (see the batched-gradient
// kernels below).
struct GradDesc {
  FwdArgs f;               // forward / recompute: geometry, coefficients, receivers, outputs
  AdjArgs b;               // adjoint: adjoint sources, gradient accumulators
  float* A; float* B;      // state ping-pong
  float* LA; float* LB;    // adjoint-state ping-pong
  float* D;                // [k][S][plane] brackets of the current segment, by global step
  char* slots;             // checkpoint slots of segments c0 .. nseg-1
  float* G1;               // [groups][plane] imaging sums
  float* GW;               // [S][nt] wavelet-gradient samples
  const int* rcell;        // [R] padded cell of each receiver
  const float* vp; const float* rho;   // the problem's model (chain rule)
  float* grad_vp;          // (nz, nx) output or null
  float* grad_w;           // shape of src_v, output or null
  float dz;
  int shared;              // wavelet shared by every shot
  long long fstride;       // S * plane
  long long slot_bytes;    // bytes per checkpoint slot (3 * fstride floats)
  int off;                 // global step of the problem's first step
  int c0;                  // segment holding that step
  int groups;              // shot groups (units = ntiles * groups)
  int pad;
};

__device__ __forceinline__ int batch_problem(const int* __restrict__ unit_start, int n_active, int u) {
  int lo = 0, hi = n_active - 1;  // last problem whose first unit is <= u
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (unit_start[mid] <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Two-step adjoint kernel: adjoint steps it and it-1 in one launch, on the
// staging geometry of the two-step forward kernel (the adjoint's stencils
// reach as far as the forward's).  Regions, relative to the tile:
//   step it:   lambda_p' rows -4 .. TZ+4, groups -1 .. GW+1;
//              lambda_v  rows -4 .. TZ+2, groups -1 .. GW (row -4 only feeds
//              the free-surface fold of step it-1)
//   step it-1: lambda_p' rows -1 .. TZ+1, groups -1 .. GW; lambda_v the tile.
// Each slot keeps its own adjoint state in registers across the phases; LDS
// holds what neighbours read: lambda_p and lambda_vz at the start of a step
// (the transposes of the free-surface mirrors read row f-1), q = a2*lambda_v
// and r = a1*lambda_p'.  The records of steps it-1 and it-2 are injected where
// the next step reads them.  Per cell the arithmetic is the one-step kernel's,
// in the same order, and the two steps' gradient sums are added in step order,
// so results equal two one-step launches bit for bit.
// ---------------------------------------------------------------------------
// PHASED: a launch of the batched gradient (arguments a_in.batch_*): the
// workgroup finds its problem, whose first step may be the launch's first
// (a.one); the single-problem instantiation compiles without either
template <int TZ, int TXT, int NTH, bool RHO, bool PHASED = false>
__global__ __launch_bounds__(NTH) void adj2_step_kernel(AdjArgs a) {
  // this kernel's tile width (columns) and the staging geometry that follows
  constexpr int TX = TXT, GW = TXT / 4, W2 = GW + 4, LW2 = GW + 6;
  constexpr int SR = TZ + 12;
  constexpr int NS = (SR * W2 + NTH - 1) / NTH;
  __shared__ float4 sLP[SR][LW2];  // lambda_p at the start of the step
  __shared__ float4 sLZ[SR][LW2];  // lambda_vz at the start of the step
  __shared__ float4 sQX[SR][LW2];  // a2 * lambda_vx on the X-region, else 0
  __shared__ float4 sQZ[SR][LW2];  // a2 * lambda_vz (folded) on the Z-region, else 0
  __shared__ float4 sR[SR][LW2];   // a1 * lambda_p' on the P-region, else 0
  __shared__ int sSrc[MAX_SPB];
  // imaging-condition sums of the tile, per step (each cell owned by one thread)
  __shared__ float4 sG[2][TZ][GW];
  __shared__ float4 sG2[RHO ? 2 : 1][RHO ? TZ : 1][RHO ? GW : 1];
  int tile = 0, group = 0;
  if constexpr (PHASED) {
    // the kernel argument carries only the batch launch: replace it by the
    // arguments of this workgroup's problem
    const int u = xcd_tile();
    const int* units = a.batch_units;
    const int n_active = a.batch_n, parity = a.batch_parity, g = a.batch_g;
    const int seg0 = a.batch_c * a.batch_k;
    if (u >= units[n_active]) return;
    const int p = batch_problem(units, n_active, u);
    const GradDesc& bd = a.batch_descs[p];
    a = bd.b;
    const float* lin = parity ? bd.LB : bd.LA;
    float* lout = parity ? bd.LA : bd.LB;
    a.LP0 = lin; a.LX0 = lin + bd.fstride; a.LZ0 = lin + 2 * bd.fstride;
    a.LP1 = lout; a.LX1 = lout + bd.fstride; a.LZ1 = lout + 2 * bd.fstride;
    a.it = g - bd.off;
    a.one = a.it == 0;
    a.dbuf = bd.D + (long long)(g - seg0) * bd.fstride;
    a.dbuf1 = a.one ? a.dbuf : a.dbuf - bd.fstride;
    a.gxbuf = a.gzbuf = a.gxbuf1 = a.gzbuf1 = nullptr;
    a.gA2 = nullptr;
    const int local = u - units[p];
    tile = local % a.ntiles;
    group = local / a.ntiles;
  }
  const Dims d = a.d;
  const int tid = threadIdx.x;
  if constexpr (!PHASED) {
    tile = xcd_tile();
    if (tile >= a.ntiles) return;
    group = blockIdx.y;
  }
  const int x0 = (tile % a.tiles_x) * TX;
  const int z0 = (tile / a.tiles_x) * TZ;
  const int s_begin = group * a.spb;
  const int s_end = min(a.S, s_begin + a.spb);
  if (s_begin >= s_end) return;  // uniform per workgroup
  const float4* __restrict__ omk1 = reinterpret_cast<const float4*>(a.coef);
  const float4* __restrict__ a1 = reinterpret_cast<const float4*>(a.coef + d.plane);
  const float4* __restrict__ omk2 = reinterpret_cast<const float4*>(a.coef + 2 * d.plane);
  const float4* __restrict__ omk3 = reinterpret_cast<const float4*>(a.coef + 3 * d.plane);
  const float4* __restrict__ a2 = reinterpret_cast<const float4*>(a.coef + 4 * d.plane);

  for (int i = tid; i < SR; i += NTH) {
    sQX[i][0] = sQX[i][LW2 - 1] = f4zero();
    sQZ[i][0] = sQZ[i][LW2 - 1] = f4zero();
    sR[i][0] = sR[i][LW2 - 1] = f4zero();
  }
  for (int i = tid; i < s_end - s_begin; i += NTH) sSrc[i] = a.src_cell[s_begin + i];

  // --- per-slot constants ---
  unsigned idx[NS], fl[NS];
  float4 c_omk1[NS], c_a1[NS], c_omk2[NS], c_omk3[NS], c_a2[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const int item = tid + n * NTH;
    const int r = item / W2, g = item % W2;
    const int gz = z0 - 6 + r, xb = x0 - 8 + 4 * g;
    bool ok;
    idx[n] = group_index(d, gz, xb, ok);
    const bool valid = item < SR * W2;
    ok = ok && valid;
    const bool p1 = valid && r >= 2 && r <= TZ + 10 && g >= 1;
    const bool v1 = valid && r >= 2 && r <= TZ + 8 && g >= 1 && g <= GW + 2;
    const bool p2 = valid && r >= 5 && r <= TZ + 7 && g >= 1 && g <= GW + 2;
    const bool tt = valid && r >= 6 && r <= TZ + 5 && g >= 2 && g <= GW + 1;
    unsigned f = 0;
    if (ok) f |= F_OK;
    if (p1) f |= F_P1;
    if (v1) f |= F_V1;
    if (p2) f |= F_P2;
    if (tt) f |= F_T;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = xb + e;
      const bool in = ok && x < d.NX;
      if (in) f |= 1u << (M_PIN + e);
      if (in && gz >= d.f + 1 && gz <= d.NZ - 3 && x >= 2 && x <= d.NX - 3) f |= 1u << (M_PREG + e);
      if (in && gz >= d.f && gz <= d.NZ - 2 && x >= 1 && x <= d.NX - 3) f |= 1u << (M_XREG + e);
      if (in && gz >= d.f && gz <= d.NZ - 3 && x >= 1 && x <= d.NX - 2) f |= 1u << (M_ZREG + e);
      // vx has NX-1 columns, vz NZ-1 rows: the 1^T stencil updates those
      if (in && x <= d.NX - 2) f |= 1u << (M_ENRG + e);
    }
    if (ok && (v1 || tt) && a.rrow[gz]) f |= F_RCV;
    if (((f >> M_PREG) & 15u) == 15u) f |= F_PFULL;
    // every cell inside the grid and in the P, X and Z regions, with the vx
    // column and vz row the 1^T stencil updates, off the free-surface rows
    if (ok && ((f >> M_PREG) & 15u) == 15u && ((f >> M_XREG) & 15u) == 15u &&
        ((f >> M_ZREG) & 15u) == 15u && ((f >> M_ENRG) & 15u) == 15u && gz <= d.NZ - 2 &&
        !(d.fs && gz >= d.f - 1 && gz <= d.f + 1))
      f |= F_AINT;
    fl[n] = f;
    c_omk1[n] = p1 ? ld4u(omk1, idx[n]) : f4zero();
    c_a1[n] = p1 ? ld4u(a1, idx[n]) : f4zero();
    c_omk2[n] = v1 ? ld4u(omk2, idx[n]) : f4zero();
    c_omk3[n] = v1 ? ld4u(omk3, idx[n]) : f4zero();
    c_a2[n] = ld4u(a2, idx[n]);
  }

  float4 rlp[NS], rlx[NS], rlz[NS];
  auto load = [&](int s) {
    const long long so = (long long)s * d.plane;
    const float4* __restrict__ LP = reinterpret_cast<const float4*>(a.LP0 + so);
    const float4* __restrict__ LX = reinterpret_cast<const float4*>(a.LX0 + so);
    const float4* __restrict__ LZ = reinterpret_cast<const float4*>(a.LZ0 + so);
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      rlp[n] = ld4u(LP, idx[n]);
      rlx[n] = ld4u(LX, idx[n]);
      rlz[n] = ld4u(LZ, idx[n]);
    }
  };
  __syncthreads();  // sSrc
  load(s_begin);

#pragma unroll
  for (int n = 0; n < NS; ++n) {
    if (fl[n] & F_T) {
      const int item = tid + n * NTH;
      const int tr = item / W2 - 6, h = item % W2 - 2;
      sG[0][tr][h] = sG[1][tr][h] = f4zero();
      if constexpr (RHO) sG2[0][tr][h] = sG2[1][tr][h] = f4zero();
    }
  }

  const int rfl = d.f - (z0 - 6);  // LDS row of the free-surface row f
  // this slot's own r (neighbours read it from LDS)
  float4 own_r[NS];
  // q = a2 * lambda_v of one slot (X / Z regions only)
  auto make_q = [&](int n, const float4& lxv, const float4& lzv, float4& qx, float4& qz) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      f4set(qx, e, f4get(c_a2[n], e) * f4get(lxv, e));
      f4set(qz, e, f4get(c_a2[n], e) * f4get(lzv, e));
    }
    if (!(fl[n] & F_AINT)) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((fl[n] >> (M_XREG + e)) & 1u)) f4set(qx, e, 0.f);
        if (!((fl[n] >> (M_ZREG + e)) & 1u)) f4set(qz, e, 0.f);
      }
    }
  };
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const int scell = sSrc[s - s_begin];
    const int ssz = scell / d.PX, ssx = scell - ssz * d.PX;
    const bool src_tile = ssz >= z0 && ssz < z0 + TZ && ssx >= x0 && ssx < x0 + TX;
    // Re-derive the region tests from fl each shot: left loop-invariant, the
    // compiler keeps each as a 64-bit lane mask in SGPRs across the loop and
    // spills them to VGPR lanes (about 90 readlanes per shot; 1.6% slower).
#pragma unroll
    for (int n = 0; n < NS; ++n) asm volatile("" : "+v"(fl[n]));
    // adjoint state of this slot (registers)
    float4 lp[NS], lx[NS], lz[NS];
    // D (and brackets) of step it for the tile slots (the others read a valid
    // group), issued before the next shot's prefetch so that waiting for them
    // does not wait for it; step it-1's are issued when that step starts
    float4 dA[NS], dB[NS], gxA[RHO ? NS : 1], gzA[RHO ? NS : 1];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const bool ok = fl[n] & F_OK;
      lp[n] = ok ? rlp[n] : f4zero();
      lx[n] = ok ? rlx[n] : f4zero();
      lz[n] = ok ? rlz[n] : f4zero();
      dA[n] = ld4u(reinterpret_cast<const float4*>(a.dbuf + so), idx[n]);
      if constexpr (!RHO) dB[n] = ld4u(reinterpret_cast<const float4*>(a.dbuf1 + so), idx[n]);
      if constexpr (RHO) {
        gxA[n] = ld4u(reinterpret_cast<const float4*>(a.gxbuf + so), idx[n]);
        gzA[n] = ld4u(reinterpret_cast<const float4*>(a.gzbuf + so), idx[n]);
      }
      const int item = tid + n * NTH;
      if (d.fs && item < SR * W2) {  // only the free-surface transposes read these
        const int r = item / W2, c = item % W2 + 1;
        sLP[r][c] = lp[n];
        sLZ[r][c] = lz[n];
      }
    }
    // the q phase below writes only LDS arrays the previous shot last read
    // before its final barrier; sLP / sLZ are read only with a free surface
    if (d.fs) __syncthreads();
    // step it: q, with the transpose of vz[f-1] = vz[f] folded into lambda_vz
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const int item = tid + n * NTH;
      if (item < SR * W2) {
        const int r = item / W2, c = item % W2 + 1;
        const int gz = z0 - 6 + r;
        float4 v = lz[n];
        if (d.fs && (gz == d.f || gz == d.f - 1)) {
          if (gz == d.f) {
            float4 up;
            if (r >= 1) {
              up = sLZ[r - 1][c];
            } else {  // row f-1 lies above the staged rows
              up = ld4u(reinterpret_cast<const float4*>(a.LZ0 + so), idx[n] - (unsigned)(d.PX >> 2));
            }
            v = make_float4(v.x + up.x, v.y + up.y, v.z + up.z, v.w + up.w);
          } else {
            v = f4zero();
          }
        }
        lz[n] = v;
        float4 qx, qz;
        make_q(n, lx[n], v, qx, qz);
        sQX[r][c] = qx;
        sQZ[r][c] = qz;
      }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int step = a.it - k;
      // the next shot's loads overlap this shot's second step
      if (k == 1 && s + 1 < s_end) load(s + 1);
      if (PHASED && k == 1 && a.one) break;  // uniform: step it was the problem's first step
      const unsigned preg_s = k == 0 ? F_P1 : F_P2;
      const unsigned vreg_s = k == 0 ? F_V1 : F_T;
      float4 dcur[NS], gxc[RHO ? NS : 1], gzc[RHO ? NS : 1];
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        if (k == 0) {
          dcur[n] = dA[n];
          if constexpr (RHO) {
            gxc[n] = gxA[n];
            gzc[n] = gzA[n];
          }
        } else {
          if constexpr (RHO) dcur[n] = ld4u(reinterpret_cast<const float4*>(a.dbuf1 + so), idx[n]);
          else dcur[n] = dB[n];
          if constexpr (RHO) {
            gxc[n] = ld4u(reinterpret_cast<const float4*>(a.gxbuf1 + so), idx[n]);
            gzc[n] = ld4u(reinterpret_cast<const float4*>(a.gzbuf1 + so), idx[n]);
          }
        }
      }
      // lambda_p' (gather of q; transpose of p[f-1] = -p[f+1]) and r = a1 lambda_p'
      // (r goes to its own array: nothing reads it until after the barrier)
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const int item = tid + n * NTH;
        if (item >= SR * W2) continue;
        const int r = item / W2, c = item % W2 + 1;
        float4 rr = f4zero();
        if (fl[n] & preg_s) {
          const int gz = z0 - 6 + r;
          // this slot's own q, recomputed from its registers rather than read back
          float4 qb, z00;
          make_q(n, lx[n], lz[n], qb, z00);
          const float4 qa = sQX[r][c - 1], qc = sQX[r][c + 1];
          const float qx[12] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w, qc.x, qc.y, qc.z, qc.w};
          const float4 zm2 = sQZ[r - 2][c], zm1 = sQZ[r - 1][c];
          const float4 zp1 = sQZ[r + 1][c];
          float4 lpv;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float az = kC1 * (f4get(z00, e) - f4get(zm1, e)) + kC2 * (f4get(zp1, e) - f4get(zm2, e));
            const float ax = kC1 * (qx[4 + e] - qx[3 + e]) + kC2 * (qx[5 + e] - qx[2 + e]);
            float acc = az;
            acc = acc + ax;
            f4set(lpv, e, f4get(lp[n], e) + acc);
          }
          if (!(fl[n] & F_AINT)) {
            if (!(fl[n] & F_OK)) {
              lpv = f4zero();
            } else if (d.fs && gz == d.f + 1) {
              const float4 lpm4 = sLP[r - 2][c];  // lambda_p at row f-1
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                // lambda_p'[f-1] = lambda_p[f-1] + c2 * q_z[f]: rows above f carry no q
                const float zero = 0.f;
                const float azm = kC1 * (zero - zero) + kC2 * (f4get(zm1, e) - zero);
                const float axm = kC1 * (zero - zero) + kC2 * (zero - zero);
                float accm = azm;
                accm = accm + axm;
                const float lpm = f4get(lpm4, e) + accm;
                f4set(lpv, e, f4get(lpv, e) - lpm);
              }
            } else if (d.fs && gz == d.f - 1) {
              lpv = f4zero();
            }
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) f4set(rr, e, f4get(c_a1[n], e) * f4get(lpv, e));
          if (!(fl[n] & F_PFULL)) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (!((fl[n] >> (M_PREG + e)) & 1u)) f4set(rr, e, 0.f);
          }
          lp[n] = lpv;  // lambda_p'
        }
        own_r[n] = rr;
        sR[r][c] = rr;
      }
      __syncthreads();
      // lambda_v, lambda_p, imaging condition, records of step-1; after the
      // first step also q of the second (the next step's inputs are this
      // slot's own outputs, except the free-surface fold below)
      const bool inject = step >= 1;
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        if (fl[n] & vreg_s) {
          const int item = tid + n * NTH;
          const int r = item / W2, c = item % W2 + 1;
          const int gz = z0 - 6 + r;
          const bool own = (fl[n] & (F_T | F_OK)) == (F_T | F_OK);
          const unsigned ti = idx[n];
          const float4 ra = sR[r][c - 1], rb = own_r[n], rc = sR[r][c + 1];
          const float rs[12] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w, rc.x, rc.y, rc.z, rc.w};
          const float4 ru = sR[r - 1][c], rd1v = sR[r + 1][c], rd2v = sR[r + 2][c];
          float4 lpo, lx4, lz4;
          if (fl[n] & F_AINT) {  // every cell in every region: no masks
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const float lpe = f4get(lp[n], e);
              f4set(lpo, e, f4get(c_omk1[n], e) * lpe);
              float lxe = f4get(lx[n], e), lze = f4get(lz[n], e);
              lxe = f4get(c_omk2[n], e) * lxe;
              lze = f4get(c_omk3[n], e) * lze;
              lxe += kC1 * (rs[5 + e] - rs[4 + e]) + kC2 * (rs[6 + e] - rs[3 + e]);
              lze += kC1 * (f4get(rd1v, e) - f4get(rb, e)) + kC2 * (f4get(rd2v, e) - f4get(ru, e));
              f4set(lx4, e, lxe);
              f4set(lz4, e, lze);
            }
            if (own) {
              float4 g = sG[k][r - 6][c - 3];
#pragma unroll
              for (int e = 0; e < 4; ++e)
                f4set(g, e, f4get(g, e) - f4get(lp[n], e) * f4get(dcur[n], e));
              sG[k][r - 6][c - 3] = g;
              if constexpr (RHO) {
                float4 g2 = sG2[k][r - 6][c - 3];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  float g2e = f4get(g2, e);
                  g2e = g2e - f4get(lz[n], e) * f4get(gzc[n], e);
                  g2e = g2e - f4get(lx[n], e) * f4get(gxc[n], e);
                  f4set(g2, e, g2e);
                }
                sG2[k][r - 6][c - 3] = g2;
              }
            }
          } else {
            lpo = lp[n];
            lx4 = lx[n];
            lz4 = lz[n];
            float4 g = f4zero(), g2 = f4zero();
            if (own) {
              g = sG[k][r - 6][c - 3];
              if constexpr (RHO) g2 = sG2[k][r - 6][c - 3];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const float lpe = f4get(lp[n], e);
              if ((fl[n] >> (M_PREG + e)) & 1u) {
                if (own) f4set(g, e, f4get(g, e) - lpe * f4get(dcur[n], e));
                f4set(lpo, e, f4get(c_omk1[n], e) * lpe);
              }
              float lxe = f4get(lx4, e), lze = f4get(lz4, e);
              if constexpr (RHO) {
                if (own) {
                  float g2e = f4get(g2, e);
                  if ((fl[n] >> (M_ZREG + e)) & 1u) g2e = g2e - lze * f4get(gzc[n], e);
                  if ((fl[n] >> (M_XREG + e)) & 1u) g2e = g2e - lxe * f4get(gxc[n], e);
                  f4set(g2, e, g2e);
                }
              }
              if ((fl[n] >> (M_XREG + e)) & 1u) lxe = f4get(c_omk2[n], e) * lxe;
              if ((fl[n] >> (M_ZREG + e)) & 1u) lze = f4get(c_omk3[n], e) * lze;
              if ((fl[n] >> (M_ENRG + e)) & 1u)
                lxe += kC1 * (rs[5 + e] - rs[4 + e]) + kC2 * (rs[6 + e] - rs[3 + e]);
              if (((fl[n] >> (M_PIN + e)) & 1u) && gz <= d.NZ - 2)
                lze += kC1 * (f4get(rd1v, e) - f4get(rb, e)) + kC2 * (f4get(rd2v, e) - f4get(ru, e));
              f4set(lx4, e, lxe);
              f4set(lz4, e, lze);
            }
            if (own) {
              sG[k][r - 6][c - 3] = g;
              if constexpr (RHO) sG2[k][r - 6][c - 3] = g2;
            }
          }
          if (src_tile && own && a.gw) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if ((int)(ti * 4u + e) == scell && ((fl[n] >> (M_PIN + e)) & 1u))
                a.gw[(long long)s * a.nt + step] = a.dt * f4get(lp[n], e);
          }
          if (inject && (fl[n] & F_RCV)) {
            const int4 rh = reinterpret_cast<const int4*>(a.rhead)[ti];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              float lpv = f4get(lpo, e), lxe = f4get(lx4, e), lze = f4get(lz4, e);
              for (int hh = e == 0 ? rh.x : (e == 1 ? rh.y : (e == 2 ? rh.z : rh.w)); hh >= 0;
                   hh = a.rnext[hh]) {
                const long long o = ((long long)s * a.nt + (step - 1)) * a.R + hh;
                if (a.g_p) lpv += a.g_p[o];
                if (a.g_vx) lxe += a.g_vx[o];
                if (a.g_vz) lze += a.g_vz[o];
              }
              f4set(lpo, e, lpv);
              f4set(lx4, e, lxe);
              f4set(lz4, e, lze);
            }
          }
          lp[n] = lpo;
          lx[n] = lx4;
          if (k == 0) {
            if (d.fs) {
              sLP[r][c] = lpo;  // read by the next step's transpose at row f+1
              sLZ[r][c] = lz4;  // read by the next step's fold at row f
            }
            // transpose of vz[f-1] = vz[f] for the next step: row f-1 gives its
            // value to row f (added below, after the barrier) and becomes 0
            if (d.fs && gz == d.f - 1) lz4 = f4zero();
            lz[n] = lz4;
            float4 qx, qz;
            make_q(n, lx4, lz4, qx, qz);
            sQX[r][c] = qx;
            sQZ[r][c] = qz;
          } else if (own) {
            st4u(reinterpret_cast<float4*>(a.LP1 + so), ti, lpo);
            st4u(reinterpret_cast<float4*>(a.LX1 + so), ti, lx4);
            st4u(reinterpret_cast<float4*>(a.LZ1 + so), ti, lz4);
          }
        }
      }
      __syncthreads();
      if (k == 0 && d.fs && rfl >= 3 && rfl <= TZ + 8) {
        // the next step's fold at row f: lambda_vz[f] += lambda_vz[f-1], then its q
#pragma unroll
        for (int n = 0; n < NS; ++n) {
          const int item = tid + n * NTH;
          if ((fl[n] & F_V1) && item / W2 == rfl) {
            const int r = rfl, c = item % W2 + 1;
            const float4 up = sLZ[r - 1][c];
            const float4 v = lz[n];
            lz[n] = make_float4(v.x + up.x, v.y + up.y, v.z + up.z, v.w + up.w);
            float4 qx, qz;
            make_q(n, lx[n], lz[n], qx, qz);
            sQZ[r][c] = qz;
          }
        }
        __syncthreads();
      }
    }
  }
  // the two steps' sums in step order: as two one-step launches add them
  float4* g1 = reinterpret_cast<float4*>(a.gA1 + (long long)group * d.plane);
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    if ((fl[n] & (F_T | F_OK)) == (F_T | F_OK)) {
      const int item = tid + n * NTH;
      const int tr = item / W2 - 6, h = item % W2 - 2;
      const float4 v = ld4u(g1, idx[n]), a0 = sG[0][tr][h], a1v = sG[1][tr][h];
      float4 w = make_float4(v.x + a0.x, v.y + a0.y, v.z + a0.z, v.w + a0.w);
      if (!(PHASED && a.one)) w = make_float4(w.x + a1v.x, w.y + a1v.y, w.z + a1v.z, w.w + a1v.w);
      st4u(g1, idx[n], w);
      if constexpr (RHO) {
        float4* g2 = reinterpret_cast<float4*>(a.gA2 + (long long)group * d.plane);
        const float4 u = ld4u(g2, idx[n]), b0 = sG2[0][tr][h], b1 = sG2[1][tr][h];
        float4 t = make_float4(u.x + b0.x, u.y + b0.y, u.z + b0.z, u.w + b0.w);
        if (!(PHASED && a.one)) t = make_float4(t.x + b1.x, t.y + b1.y, t.z + b1.z, t.w + b1.w);
        st4u(g2, idx[n], t);
      }
    }
  }
}


// ---------------------------------------------------------------------------
// Two adjoint steps (it, it-1) by z-marching wavefronts (launch choice
// march = 3; no density gradient).  The per-cell arithmetic
// is adj2_step_kernel's, in the same order, and the imaging sums are added in
// the same order, so the results equal it bit for bit.
//
// A wavefront holds one shot on a 128-column window (2 cells per lane) and
// walks a z-segment of rows top to bottom.  Iteration i finishes row i of
// the second step and works on five rows in a pipeline:
//   row i+6: the adjoint state lambda (step it+1), q0 = a2 * lambda_v;
//   row i+5: lambda_p' of step it (gather of q0), r1 = a1 * lambda_p',
//            lambda_p after step it, the imaging term of step it;
//   row i+3: lambda_v after step it (gather of r1), q1 = a2 * lambda_v;
//   row i+2: lambda_p' of step it-1, r2, lambda_p after it-1 (stored), the
//            imaging term of step it-1;
//   row i:   lambda_v after step it-1 (stored).
// Each stage reads the rows above and below it from a four-row window kept in
// registers, and the neighbouring columns from the adjacent lanes (DPP
// wavefront shifts), so the stencils need no barrier.  The window is 6
// columns wider than the 116 columns a wavefront owns on each side, and a
// segment's pipeline starts 6 rows above its first row.  Rows a stage works on
// outside what the segment's own rows need are computed and not used.
// The NW wavefronts of a workgroup take the shots of one shot group on the
// same window and segment: they share the coefficient rows, which the
// workgroup stages in an LDS ring one 4-row block ahead.  Each wavefront loads
// its shot's rows one iteration ahead.  Every 4 rows the wavefronts write their
// imaging terms to LDS; after a barrier the workgroup sums them over its shots
// in shot order and adds them into the group's gradient plane.
// Blocks away from the grid's edge rows, the segment's first and last rows,
// receivers and the shot's source run without any row test (G = false).
// ---------------------------------------------------------------------------
constexpr int AM_OWN = 116;   // columns a wavefront owns (64 lanes x 2 cells, less 6 per side)
constexpr int AM_SLOTS = 8;   // LDS row slots of the imaging terms: two 4-row blocks
constexpr int AM_CROWS = 16;  // LDS ring rows of the coefficients
#ifndef GBW_AM_TARGET_WGS
#define GBW_AM_TARGET_WGS 256
#endif
constexpr int AM_TARGET_WGS = GBW_AM_TARGET_WGS;  // workgroups to aim for (MI355X: 256 CUs)
#ifndef GBW_AM_PF
#define GBW_AM_PF 2
#endif
constexpr int AM_PF = GBW_AM_PF;  // iterations a wavefront loads its shot's rows ahead (1, 2, 4)
static_assert(AM_PF == 1 || AM_PF == 2 || AM_PF == 4, "the 4-row unroll needs AM_PF | 4");

__device__ __forceinline__ float2 f2zero() { return make_float2(0.f, 0.f); }
// float2 at byte offset `lane_off` of row `row` of a plane (uniform row base,
// 32-bit lane offset)
__device__ __forceinline__ float2 row_ld2(const float* plane, int row, int PX, unsigned lane_off) {
  const char* rb = reinterpret_cast<const char*>(plane) + (long long)row * PX * 4;
  return *reinterpret_cast<const float2*>(rb + lane_off);
}
__device__ __forceinline__ void row_st2(float* plane, int row, int PX, unsigned lane_off,
                                        const float2& v) {
  char* rb = reinterpret_cast<char*>(plane) + (long long)row * PX * 4;
  *reinterpret_cast<float2*>(rb + lane_off) = v;
}

// CM: the window reaches a column outside the P, X or Z region (first and
// last strips): per-cell column masks.  FS: free surface (the mirror
// transposes at rows f-1 .. f+1)
template <int NW, bool CM, bool FS>
__device__ __forceinline__ void adj2m_body(const AdjArgs& a, int group, int seg, int strip,
                                           float2 (*sT)[NW][AM_SLOTS][64],
                                           float2 (*sC)[AM_CROWS][64]) {
  const Dims d = a.d;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s_begin = group * a.spb;
  const int ns = min(a.S, s_begin + a.spb) - s_begin;
  const bool active = w < ns;  // uniform per wavefront
  const int s = s_begin + (active ? w : 0);
  const int zs = seg == 0 ? 0 : a.am_seg0_h + (seg - 1) * a.am_seg_h;
  const int ze = min(d.NZ, seg == 0 ? a.am_seg0_h : zs + a.am_seg_h);
  const int xs = strip * AM_OWN - 6;
  const int x0 = xs + 2 * lane;
  const bool lane_ok = x0 >= 0 && x0 <= d.PX - 2;
  const unsigned xo4 = (unsigned)min(max(x0, 0), d.PX - 2) * 4u;  // lane byte offset in a row
  const bool own_lane = lane >= 3 && lane < 3 + AM_OWN / 2 && x0 < d.NX;
  // column tests of the lane's two cells
  bool cin[2], cp[2], cx[2], cz[2], ce[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int x = x0 + e;
    cin[e] = lane_ok && x < d.NX;
    cp[e] = cin[e] && x >= 2 && x <= d.NX - 3;
    cx[e] = cin[e] && x >= 1 && x <= d.NX - 3;
    cz[e] = cin[e] && x >= 1 && x <= d.NX - 2;
    ce[e] = cin[e] && x <= d.NX - 2;
  }
  const long long so = (long long)s * d.plane;
  const float* LP0 = a.LP0 + so;
  const float* LX0 = a.LX0 + so;
  const float* LZ0 = a.LZ0 + so;
  float* LP1 = a.LP1 + so;
  float* LX1 = a.LX1 + so;
  float* LZ1 = a.LZ1 + so;
  const float* D0 = a.dbuf + so;
  const float* D1 = a.dbuf1 + so;
  const int scell = a.src_cell[s];
  const int ssz = scell / d.PX, ssx = scell - ssz * d.PX;
  const int it = a.it;
  const float2* sCl = &sC[0][0][lane];  // this lane's column in the coefficient ring
  auto row_in = [&](int r) { return r >= 0 && r < d.NZ; };
  auto row_p = [&](int r) { return r >= d.f + 1 && r <= d.NZ - 3; };
  auto row_x = [&](int r) { return r >= d.f && r <= d.NZ - 2; };
  auto row_z = [&](int r) { return r >= d.f && r <= d.NZ - 3; };
  // receivers' adjoint sources of record step `rec` at row r, added to the
  // lane's cells of p and / or v (same list order as the tile kernels)
  auto inject = [&](int r, int rec, float2* lp, float2* lx, float2* lz) {
    if (!lane_ok) return;
    const int2 rh = *reinterpret_cast<const int2*>(
        reinterpret_cast<const char*>(a.rhead) + (long long)r * d.PX * 4 + xo4);
    for (int hh = rh.x; hh >= 0; hh = a.rnext[hh]) {
      const long long o = ((long long)s * a.nt + rec) * a.R + hh;
      if (lp && a.g_p) lp->x = lp->x + a.g_p[o];
      if (lx && a.g_vx) lx->x = lx->x + a.g_vx[o];
      if (lz && a.g_vz) lz->x = lz->x + a.g_vz[o];
    }
    for (int hh = rh.y; hh >= 0; hh = a.rnext[hh]) {
      const long long o = ((long long)s * a.nt + rec) * a.R + hh;
      if (lp && a.g_p) lp->y = lp->y + a.g_p[o];
      if (lx && a.g_vx) lx->y = lx->y + a.g_vx[o];
      if (lz && a.g_vz) lz->y = lz->y + a.g_vz[o];
    }
  };
  // The stage helpers take the row's coefficients through a getter C(plane)
  // (plane: 0 omk1, 1 a1, 2 omk2, 3 omk3, 4 a2): an LDS ring read where the
  // helper needs it, or a value read earlier (see the iteration).
  // q = a2 * lambda_v of row r (X / Z regions only); G: test the row
  auto make_q = [&](auto G, int r, auto C, const float2& lx, const float2& lz,
                    float2& qx, float2& qz) {
    const float2 c2 = C(4);
    qx = make_float2(c2.x * lx.x, c2.y * lx.y);
    qz = make_float2(c2.x * lz.x, c2.y * lz.y);
    if constexpr (decltype(G)::value) {
      if (!row_x(r)) qx = f2zero();
      if (!row_z(r)) qz = f2zero();
    }
    if constexpr (CM) {
      if (!cx[0]) qx.x = 0.f;
      if (!cx[1]) qx.y = 0.f;
      if (!cz[0]) qz.x = 0.f;
      if (!cz[1]) qz.y = 0.f;
    }
  };
  // lambda_p' = lambda_p + gather of q (rows r-2..r+1 of qz, columns of qx),
  // then r = a1 * lambda_p' on the P region (pr: cells in it) and lambda_p
  // after the step, omk1 * lambda_p' on the P region
  // With a free surface, the transpose of p[f-1] = -p[f+1]: lambda_p' of row
  // f-1 (kept in *sfm) is taken from row f+1 and row f-1 becomes 0 (rows above
  // f carry no q, so row f-1's gather is the reference's closed form)
  auto lam_p = [&](auto G, int r, auto C, const float2& lp, const float2& qx,
                   const float2& zm2, const float2& zm1, const float2& z00, const float2& zp1,
                   float2& lpv, float2& rr, float2& lpo, bool* pr, float2* sfm) {
    const float ql0 = lane_prev(qx.x), ql1 = lane_prev(qx.y), qr0 = lane_next(qx.x);
    {
      const float az = kC1 * (z00.x - zm1.x) + kC2 * (zp1.x - zm2.x);
      const float ax = kC1 * (qx.x - ql1) + kC2 * (qx.y - ql0);
      float acc = az;
      acc = acc + ax;
      lpv.x = lp.x + acc;
    }
    {
      const float az = kC1 * (z00.y - zm1.y) + kC2 * (zp1.y - zm2.y);
      const float ax = kC1 * (qx.y - qx.x) + kC2 * (qr0 - ql1);
      float acc = az;
      acc = acc + ax;
      lpv.y = lp.y + acc;
    }
    bool rp = true;
    if constexpr (decltype(G)::value) {
      rp = row_p(r);
      if (!row_in(r)) lpv = f2zero();
    }
    if constexpr (CM) {
      if (!lane_ok) lpv = f2zero();
    }
    if constexpr (FS && decltype(G)::value) {
      if (r == d.f - 1) {
        *sfm = lpv;
        lpv = f2zero();
      } else if (r == d.f + 1) {
        lpv.x = lpv.x - sfm->x;
        lpv.y = lpv.y - sfm->y;
      }
    }
    const float2 c1 = C(1), o1 = C(0);
    rr = make_float2(c1.x * lpv.x, c1.y * lpv.y);
    lpo = make_float2(o1.x * lpv.x, o1.y * lpv.y);
    pr[0] = rp && (!CM || cp[0]);
    pr[1] = rp && (!CM || cp[1]);
    if constexpr (CM || decltype(G)::value) {
      if (!pr[0]) { rr.x = 0.f; lpo.x = lpv.x; }
      if (!pr[1]) { rr.y = 0.f; lpo.y = lpv.y; }
    }
  };
  // lambda_v after the step: omk * lambda_v + gather of r (rows r-1..r+2)
  auto lam_v = [&](auto G, int r, auto C, float2& lx, float2& lz, const float2& rm1,
                   const float2& r00, const float2& rp1, const float2& rp2) {
    constexpr bool GG = decltype(G)::value;
    const float rl1 = lane_prev(r00.y), rr0 = lane_next(r00.x), rr1 = lane_next(r00.y);
    const float2 o2 = C(2), o3 = C(3);
    bool rx = true, rz = true, rin = true, ml = true;
    if constexpr (GG) {
      rx = row_x(r); rz = row_z(r); rin = row_in(r); ml = rin && r <= d.NZ - 2;
    }
    const bool mx0 = rx && (!CM || cx[0]), mx1 = rx && (!CM || cx[1]);
    const bool mz0 = rz && (!CM || cz[0]), mz1 = rz && (!CM || cz[1]);
    const bool me0 = rin && (!CM || ce[0]), me1 = rin && (!CM || ce[1]);
    const bool mi0 = ml && (!CM || cin[0]), mi1 = ml && (!CM || cin[1]);
    float lx0 = lx.x, lz0 = lz.x, lx1 = lx.y, lz1 = lz.y;
    if (mx0) lx0 = o2.x * lx0;
    if (mz0) lz0 = o3.x * lz0;
    if (me0) lx0 += kC1 * (r00.y - r00.x) + kC2 * (rr0 - rl1);
    if (mi0) lz0 += kC1 * (rp1.x - r00.x) + kC2 * (rp2.x - rm1.x);
    if (mx1) lx1 = o2.y * lx1;
    if (mz1) lz1 = o3.y * lz1;
    if (me1) lx1 += kC1 * (rr0 - r00.y) + kC2 * (rr1 - r00.x);
    if (mi1) lz1 += kC1 * (rp1.y - r00.y) + kC2 * (rp2.y - rm1.y);
    lx = make_float2(lx0, lx1);
    lz = make_float2(lz0, lz1);
  };

  // free surface: lambda_p' of row f-1 in each step, lambda_vz of row f-1
  // after the first step (folded into row f for the second)
  float2 sfm0 = f2zero(), sfm1 = f2zero(), szf = f2zero();
  // four-row register windows, slot = row & 3
  float2 WLP[4], WLX[4], WLZ[4], WQX0[4], WQZ0[4], WR1[4], WLP1[4], WT1[4];
  float2 WLX1[4], WLZ1[4], WQX1[4], WQZ1[4], WR2[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    WLP[k] = WLX[k] = WLZ[k] = WQX0[k] = WQZ0[k] = WR1[k] = WLP1[k] = WT1[k] = f2zero();
    WLX1[k] = WLZ1[k] = WQX1[k] = WQZ1[k] = WR2[k] = f2zero();
  }
  // this shot's rows, loaded AM_PF iterations ahead: lambda of row i+6, D^it
  // of row i+5, D^(it-1) of row i+2 (rows clamped into what the segment needs)
  const int ja = max(0, zs - 6), jb = min(d.NZ, ze + 6) - 1;
  float2 pf_lp[AM_PF], pf_lx[AM_PF], pf_lz[AM_PF], pf_d0[AM_PF], pf_d1[AM_PF];
  auto prefetch = [&](auto K, int i) {
    constexpr int k = decltype(K)::value;
    const int j = min(max(i + 6, ja), jb);
    pf_lp[k] = row_ld2(LP0, j, d.PX, xo4);
    pf_lx[k] = row_ld2(LX0, j, d.PX, xo4);
    pf_lz[k] = row_ld2(LZ0, j, d.PX, xo4);
    pf_d0[k] = row_ld2(D0, min(max(i + 5, zs), ze - 1), d.PX, xo4);
    pf_d1[k] = row_ld2(D1, min(max(i + 2, zs), ze - 1), d.PX, xo4);
  };

  auto iter = [&](auto U, auto G, int base) {
    constexpr int uu = decltype(U)::value;
    constexpr bool GG = decltype(G)::value;
    const int i = base + uu;
    constexpr int pk = uu % AM_PF;
    const float2 lp_in = pf_lp[pk], lx_in = pf_lx[pk], lz_in = pf_lz[pk], d0 = pf_d0[pk],
                 d1 = pf_d1[pk];
    prefetch(std::integral_constant<int, pk>{}, i + AM_PF);
    // the rows' coefficients: without a free surface the ten values the
    // stages use are read from the LDS ring up front, so the reads' latencies
    // overlap (adjoint 426 -> 422 us per launch, 64 shots of 390 x 1740); the
    // free-surface instance has no registers to spare and reads each where its
    // stage needs it
    constexpr int CP = AM_CROWS * 64;
    auto ring = [&](int row) { return sCl + (row & (AM_CROWS - 1)) * 64; };
    float2 cpre[10] = {};
    if constexpr (!FS) {
      const float2 *crj = ring(i + 6), *crb = ring(i + 5), *crc = ring(i + 3), *crd = ring(i + 2),
                   *cri = ring(i);
      cpre[0] = crj[4 * CP]; cpre[1] = crb[CP]; cpre[2] = crb[0]; cpre[3] = crc[2 * CP];
      cpre[4] = crc[3 * CP]; cpre[5] = crc[4 * CP]; cpre[6] = crd[CP]; cpre[7] = crd[0];
      cpre[8] = cri[2 * CP]; cpre[9] = cri[3 * CP];
    }
    // getter of row `row`'s planes: v[plane] read up front, or the ring
    const float2 z = f2zero();
    auto getter = [&](int row, float2 v0, float2 v1, float2 v2, float2 v3, float2 v4) {
      return [&, row, v0, v1, v2, v3, v4](int plane) -> float2 {
        if constexpr (FS) {
          return ring(row)[plane * CP];
        } else {
          return plane == 0 ? v0 : (plane == 1 ? v1 : (plane == 2 ? v2 : (plane == 3 ? v3 : v4)));
        }
      };
    };
    // --- row j = i+6: lambda of step it+1, q0 ---
    {
      constexpr int sj = (uu + 6) & 3;
      const int j = i + 6;
      float2 lp = lp_in, lx = lx_in, lz = lz_in;
      if constexpr (GG) {
        if (!row_in(j)) lp = lx = lz = f2zero();
      }
      if constexpr (CM) {
        if (!lane_ok) lp = lx = lz = f2zero();
      }
      if constexpr (FS && GG) {
        if (j == d.f) {  // transpose of vz[f-1] = vz[f]: row f-1 joins row f
          const float2 up = WLZ[(uu + 5) & 3];
          lz = make_float2(lz.x + up.x, lz.y + up.y);
          WLZ[(uu + 5) & 3] = f2zero();
        }
      }
      WLP[sj] = lp; WLX[sj] = lx; WLZ[sj] = lz;
      make_q(G, j, getter(i + 6, z, z, z, z, cpre[0]), lx, lz, WQX0[sj], WQZ0[sj]);
    }
    // --- row b = i+5: lambda_p' of step it, r1, lambda_p after it, imaging ---
    {
      constexpr int sb = (uu + 5) & 3;
      const int b = i + 5;
      float2 lpv, lpo;
      bool pr[2];
      lam_p(G, b, getter(i + 5, cpre[2], cpre[1], z, z, z), WLP[sb], WQX0[sb], WQZ0[(uu + 3) & 3],
            WQZ0[(uu + 4) & 3], WQZ0[sb], WQZ0[(uu + 6) & 3], lpv, WR1[sb], lpo, pr, &sfm0);
      if (!GG || (b >= zs && b < ze)) {
        float2 t1 = f2zero();
        if (pr[0]) t1.x = lpv.x * d0.x;
        if (pr[1]) t1.y = lpv.y * d0.y;
        WT1[sb] = t1;
      }
      if constexpr (GG) {
        if (it >= 1 && a.R > 0 && row_in(b) && a.rrow[b]) inject(b, it - 1, &lpo, nullptr, nullptr);
        if (a.gw && b == ssz && b >= zs && b < ze && own_lane) {
          if (x0 == ssx && cin[0]) a.gw[(long long)s * a.nt + it] = a.dt * lpv.x;
          if (x0 + 1 == ssx && cin[1]) a.gw[(long long)s * a.nt + it] = a.dt * lpv.y;
        }
      }
      WLP1[sb] = lpo;
    }
    // --- row c = i+3: lambda_v after step it, q1 ---
    {
      constexpr int sc = (uu + 3) & 3;
      const int c = i + 3;
      const auto cc = getter(i + 3, z, z, cpre[3], cpre[4], cpre[5]);
      float2 lx = WLX[sc], lz = WLZ[sc];
      lam_v(G, c, cc, lx, lz, WR1[(uu + 2) & 3], WR1[sc], WR1[uu & 3], WR1[(uu + 1) & 3]);
      if constexpr (GG) {
        if (it >= 1 && a.R > 0 && row_in(c) && a.rrow[c]) inject(c, it - 1, nullptr, &lx, &lz);
        // the second step's transpose of vz[f-1] = vz[f]
        if constexpr (FS) {
          if (c == d.f - 1) {
            szf = lz;
            lz = f2zero();
          } else if (c == d.f) {
            lz = make_float2(lz.x + szf.x, lz.y + szf.y);
          }
        }
      }
      WLX1[sc] = lx; WLZ1[sc] = lz;
      make_q(G, c, cc, lx, lz, WQX1[sc], WQZ1[sc]);
    }
    // --- row dr = i+2: lambda_p' of step it-1, r2, lambda_p (stored), imaging ---
    {
      constexpr int sd = (uu + 2) & 3;
      const int dr = i + 2;
      float2 lpv, lpo, r2;
      bool pr[2];
      lam_p(G, dr, getter(i + 2, cpre[7], cpre[6], z, z, z), WLP1[sd], WQX1[sd], WQZ1[uu & 3],
            WQZ1[(uu + 1) & 3], WQZ1[sd], WQZ1[(uu + 3) & 3], lpv, r2, lpo, pr, &sfm1);
      WR2[sd] = r2;
      if (!GG || (dr >= zs && dr < ze)) {
        if constexpr (GG) {
          if (it - 1 >= 1 && a.R > 0 && a.rrow[dr]) inject(dr, it - 2, &lpo, nullptr, nullptr);
        }
        if (own_lane) row_st2(LP1, dr, d.PX, xo4, lpo);
        float2 t2 = f2zero();
        if (pr[0]) t2.x = lpv.x * d1.x;
        if (pr[1]) t2.y = lpv.y * d1.y;
        sT[0][w][dr & (AM_SLOTS - 1)][lane] = WT1[sd];
        sT[1][w][dr & (AM_SLOTS - 1)][lane] = t2;
        if constexpr (GG) {
          if (a.gw && dr == ssz && own_lane) {
            if (x0 == ssx && cin[0]) a.gw[(long long)s * a.nt + it - 1] = a.dt * lpv.x;
            if (x0 + 1 == ssx && cin[1]) a.gw[(long long)s * a.nt + it - 1] = a.dt * lpv.y;
          }
        }
      }
    }
    // --- row i: lambda_v after step it-1 (stored) ---
    {
      constexpr int se = uu & 3;
      if (!GG || (i >= zs && i < ze)) {
        float2 lx = WLX1[se], lz = WLZ1[se];
        lam_v(G, i, getter(i, z, z, cpre[8], cpre[9], z), lx, lz, WR2[(uu + 3) & 3], WR2[se],
              WR2[(uu + 1) & 3], WR2[(uu + 2) & 3]);
        if constexpr (GG) {
          if (it - 1 >= 1 && a.R > 0 && a.rrow[i]) inject(i, it - 2, nullptr, &lx, &lz);
        }
        if (own_lane) {
          row_st2(LX1, i, d.PX, xo4, lx);
          row_st2(LZ1, i, d.PX, xo4, lz);
        }
      }
    }
  };

  // coefficient ring: per block 5 planes x 4 rows, one row-plane per wavefront
  // at a time (uniform row base, lane offset)
  constexpr int CRP = 20, CPW = (CRP + NW - 1) / NW;
  float2 cpf[CPW];
  auto coef_load = [&](int r0) {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      const int rp = w + k * NW;  // plane * 4 + row
      if (rp < CRP) {
        const int r = min(max(r0 + (rp & 3), 0), d.NZ - 1);
        cpf[k] = row_ld2(a.coef + (rp >> 2) * d.plane, r, d.PX, xo4);
      }
    }
  };
  auto coef_store = [&](int r0) {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      const int rp = w + k * NW;
      if (rp < CRP) sC[rp >> 2][(r0 + (rp & 3)) & (AM_CROWS - 1)][lane] = cpf[k];
    }
  };
  const int i0 = zs - 12;
  // rows i0 .. i0+11 for the first block
  coef_load(i0); coef_store(i0);
  coef_load(i0 + 4); coef_store(i0 + 4);
  coef_load(i0 + 8); coef_store(i0 + 8);
  __syncthreads();
  if (active) {
    prefetch(std::integral_constant<int, 0>{}, i0);
    if constexpr (AM_PF > 1) prefetch(std::integral_constant<int, 1 % AM_PF>{}, i0 + 1);
    if constexpr (AM_PF > 2) {
      prefetch(std::integral_constant<int, 2 % AM_PF>{}, i0 + 2);
      prefetch(std::integral_constant<int, 3 % AM_PF>{}, i0 + 3);
    }
  }

  constexpr int RED = 4 * 128 / (NW * 64);  // reduced cells per thread per block
  float* __restrict__ g1 = a.gA1 + (long long)group * d.plane;
  for (int base = i0; base < ze; base += 4) {
    // coefficient rows of the next block (base+12..base+15; this block uses
    // base..base+9, the ring rows base+10, base+11 are already there)
    coef_load(base + 12);
    // the gradient cells this block completes (rows base+2..base+5), loaded early
    float gv[RED];
#pragma unroll
    for (int k = 0; k < RED; ++k) {
      const int q = threadIdx.x + k * NW * 64;
      const int row = base + 2 + (q >> 7), cl = q & 127, col = xs + cl;
      const bool ok = row >= zs && row < ze && cl >= 6 && cl < 6 + AM_OWN && col < d.NX;
      gv[k] = ok ? *reinterpret_cast<const float*>(reinterpret_cast<const char*>(g1) +
                                                     (long long)row * d.PX * 4 + (unsigned)col * 4u)
                 : 0.f;
    }
    if (active) {
      // rows this block's stages touch: base .. base+9
      const int rr = base + lane;
      const bool rcv = lane < 10 && rr >= 0 && rr < d.NZ && a.rrow[rr];
      const bool plain = base >= zs && base + 8 < ze && base >= d.f + 1 && base + 9 <= d.NZ - 3 &&
                         __ballot(rcv) == 0 && !(ssz >= base && ssz <= base + 9);
      if (plain) {
        iter(std::integral_constant<int, 0>{}, std::false_type{}, base);
        iter(std::integral_constant<int, 1>{}, std::false_type{}, base);
        iter(std::integral_constant<int, 2>{}, std::false_type{}, base);
        iter(std::integral_constant<int, 3>{}, std::false_type{}, base);
      } else {
        iter(std::integral_constant<int, 0>{}, std::true_type{}, base);
        iter(std::integral_constant<int, 1>{}, std::true_type{}, base);
        iter(std::integral_constant<int, 2>{}, std::true_type{}, base);
        iter(std::integral_constant<int, 3>{}, std::true_type{}, base);
      }
    }
    coef_store(base + 12);
    __syncthreads();
    // sums over the group's shots in shot order, per step, then step it's sum
    // and step it-1's added to the plane (as adj2_step_kernel's flush)
#pragma unroll
    for (int k = 0; k < RED; ++k) {
      const int q = threadIdx.x + k * NW * 64;
      const int row = base + 2 + (q >> 7), cl = q & 127, col = xs + cl;
      if (!(row >= zs && row < ze && cl >= 6 && cl < 6 + AM_OWN && col < d.NX)) continue;
      const int slot = row & (AM_SLOTS - 1);
      const float* t0 = reinterpret_cast<const float*>(&sT[0][0][slot][0]) + cl;
      const float* t1 = reinterpret_cast<const float*>(&sT[1][0][slot][0]) + cl;
      float s0 = 0.f, s1 = 0.f;
      for (int ww = 0; ww < ns; ++ww) {
        s0 = s0 - t0[ww * AM_SLOTS * 128];
        s1 = s1 - t1[ww * AM_SLOTS * 128];
      }
      float v = gv[k] + s0;
      v = v + s1;
      *reinterpret_cast<float*>(reinterpret_cast<char*>(g1) + (long long)row * d.PX * 4 +
                                (unsigned)col * 4u) = v;
    }
  }
}

template <int NW, bool FS>
__global__ __launch_bounds__(NW * 64) void adj2m_kernel(AdjArgs a) {
  __shared__ float2 sT[2][NW][AM_SLOTS][64];  // imaging terms per step, shot, row slot, lane
  __shared__ float2 sC[5][AM_CROWS][64];      // coefficient ring: plane, row slot, lane
  const int groups = (a.S + a.spb - 1) / a.spb;
  const int u = xcd_tile();
  if (u >= a.ntiles * groups) return;  // uniform per workgroup
  const int group = u % groups;
  const int seg = (u / groups) % a.am_segs;
  const int strip = u / groups / a.am_segs;
  const int xs = strip * AM_OWN - 6;
  if (xs >= 2 && xs + 127 <= a.d.NX - 3)
    adj2m_body<NW, false, FS>(a, group, seg, strip, sT, sC);
  else
    adj2m_body<NW, true, FS>(a, group, seg, strip, sT, sC);
}

// ---------------------------------------------------------------------------
// Two forward steps (it, it+1) by z-marching wavefronts (launch choice
// fwd_march = 2; no free surface, no density-gradient brackets).  The window,
// segment and shot-group geometry is adj2m_kernel's: one wavefront holds one
// shot on a 128-column window (2 cells per lane, 116 owned) and walks a z
// segment top to bottom; the NW wavefronts of a workgroup are the shots of one
// shot group and share the coefficient rows through the LDS ring.  Iteration
// i finishes row i of the second step and works on five rows at once:
//   row i+6: vz of the input state (loaded; p and vx of row i+5 with it);
//   row i+5: p' of step it (vx neighbours by DPP, vz rows i+3 .. i+6);
//   row i+3: vx', vz' of step it (p' rows i+2 .. i+5), its receivers, energy;
//   row i+2: p'' of step it+1 (vx', vz' rows i .. i+3);
//   row i:   vx'', vz'' of step it+1 (p'' rows i-1 .. i+2); row i stored.
// The per-cell arithmetic is fwd2_step_kernel's, in the same order, so the
// state, the seismograms and the recompute brackets equal it bit for bit.
// Energy: each wavefront adds its shot's squares of the two steps, and every
// 4 rows the workgroup sums them over its shots in shot order and adds them to
// the group's partial sums (the same sums as the tiled kernel, associated per
// shot first; the maps agree with it to rounding).
// ---------------------------------------------------------------------------
constexpr int FM_ESLOTS = 8;  // LDS row slots of the energy squares: two 4-row blocks

template <int NW, bool CM, bool RECOMP>
__device__ __forceinline__ void fwd2m_body(const FwdArgs& a, int group, int seg, int strip,
                                           float2 (*sE)[NW][FM_ESLOTS][64],
                                           float2 (*sC)[AM_CROWS][64]) {
  const Dims d = a.d;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s_begin = group * a.spb;
  const int ns = min(a.S, s_begin + a.spb) - s_begin;
  const bool active = w < ns;  // uniform per wavefront
  const int s = s_begin + (active ? w : 0);
  const int zs = seg * a.seg_h, ze = min(d.NZ, zs + a.seg_h);
  const int xs = strip * AM_OWN - 6;
  const int x0 = xs + 2 * lane;
  const bool lane_ok = x0 >= 0 && x0 <= d.PX - 2;
  const unsigned xo4 = (unsigned)min(max(x0, 0), d.PX - 2) * 4u;  // lane byte offset in a row
  const bool own_lane = lane >= 3 && lane < 3 + AM_OWN / 2 && x0 < d.NX;
  // column tests of the lane's two cells: in the grid; P, X, Z regions
  bool cin[2], cp[2], cx[2], cz[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int x = x0 + e;
    cin[e] = lane_ok && x < d.NX;
    cp[e] = cin[e] && x >= 2 && x <= d.NX - 3;
    cx[e] = cin[e] && x >= 1 && x <= d.NX - 3;
    cz[e] = cin[e] && x >= 1 && x <= d.NX - 2;
  }
  const long long so = (long long)s * d.plane;
  const float* P0 = a.P0 + so;
  const float* VX0 = a.VX0 + so;
  const float* VZ0 = a.VZ0 + so;
  float* P1 = a.P1 + so;
  float* VX1 = a.VX1 + so;
  float* VZ1 = a.VZ1 + so;
  float* D0 = RECOMP ? a.dbuf + so : nullptr;
  float* D1 = RECOMP ? a.dbuf + a.bstride + so : nullptr;
  const int scell = a.src_cell[s];
  const int ssz = scell / d.PX, ssx = scell - ssz * d.PX;
  const float* wv = a.wav + (long long)s * a.wav_stride + a.it;
  const float supd0 = a.dt * wv[0], supd1 = a.dt * wv[1];
  const int it = a.it;
  const bool rec = !RECOMP && a.rec_p != nullptr;
  const float2* sCl = &sC[0][0][lane];  // this lane's column in the coefficient ring
  auto row_in = [&](int r) { return r >= 0 && r < d.NZ; };
  auto row_p = [&](int r) { return r >= d.f + 1 && r <= d.NZ - 3; };
  auto row_x = [&](int r) { return r >= d.f && r <= d.NZ - 2; };
  auto row_z = [&](int r) { return r >= d.f && r <= d.NZ - 3; };
  auto seg_row = [&](int r) { return r >= zs && r < ze; };
  // the receivers of row r record p, vx, vz of step `step` (same list order as
  // the tile kernels; one lane per cell)
  auto record = [&](int r, int step, const float2& pv, const float2& vx, const float2& vz) {
    const int2 rh = *reinterpret_cast<const int2*>(
        reinterpret_cast<const char*>(a.rhead) + (long long)r * d.PX * 4 + xo4);
    const long long ob = ((long long)s * a.nt + step) * a.R;
    for (int hh = rh.x; hh >= 0; hh = a.rnext[hh]) {
      a.rec_p[ob + hh] = pv.x;
      a.rec_vx[ob + hh] = vx.x;
      a.rec_vz[ob + hh] = vz.x;
    }
    for (int hh = rh.y; hh >= 0; hh = a.rnext[hh]) {
      a.rec_p[ob + hh] = pv.y;
      a.rec_vx[ob + hh] = vx.y;
      a.rec_vz[ob + hh] = vz.y;
    }
  };
  // p-update of row r (acoustic_2d.py:213-237): vx of the row (neighbours by
  // DPP), vz of rows r-2 .. r+1, the old p; cells outside the P region keep p,
  // rows outside the grid are zero.  Returns p'; br: the bracket (0 off the region)
  auto p_upd = [&](auto G, int r, const float2* cr, const float2& p, const float2& vx,
                   const float2& zm2, const float2& zm1, const float2& z00, const float2& zp1,
                   float2& br) {
    const float vl0 = lane_prev(vx.x), vl1 = lane_prev(vx.y), vr0 = lane_next(vx.x);
    const float2 o1 = cr[0], c1 = cr[1 * AM_CROWS * 64];
    float2 pn, b;
    {
      float s1 = vx.x - vl1;
      s1 = s1 + z00.x;
      s1 = s1 - zm1.x;
      float s2 = vx.y - vl0;
      s2 = s2 + zp1.x;
      s2 = s2 - zm2.x;
      const float t1 = kC1 * s1;
      const float t2 = kC2 * s2;
      b.x = t1 + t2;
      const float u = o1.x * p.x;
      pn.x = u - c1.x * b.x;
    }
    {
      float s1 = vx.y - vx.x;
      s1 = s1 + z00.y;
      s1 = s1 - zm1.y;
      float s2 = vr0 - vl1;
      s2 = s2 + zp1.y;
      s2 = s2 - zm2.y;
      const float t1 = kC1 * s1;
      const float t2 = kC2 * s2;
      b.y = t1 + t2;
      const float u = o1.y * p.y;
      pn.y = u - c1.y * b.y;
    }
    bool rp = true;
    if constexpr (decltype(G)::value) rp = row_p(r);
    const bool m0 = rp && (!CM || cp[0]), m1 = rp && (!CM || cp[1]);
    if constexpr (CM || decltype(G)::value) {
      if (!m0) { pn.x = p.x; b.x = 0.f; }
      if (!m1) { pn.y = p.y; b.y = 0.f; }
    }
    if constexpr (decltype(G)::value) {
      if (!row_in(r)) pn = f2zero();
    }
    br = b;
    return pn;
  };
  // v-update of row r (acoustic_2d.py:243-265) from p' of rows r-1 .. r+2 (x
  // neighbours of row r by DPP); cells outside the X / Z regions keep v
  auto v_upd = [&](auto G, int r, const float2* cr, const float2& pu, const float2& pc,
                   const float2& pd1, const float2& pd2, float2& vx, float2& vz) {
    const float pl1 = lane_prev(pc.y), pr0 = lane_next(pc.x), pr1 = lane_next(pc.y);
    const float2 o2 = cr[2 * AM_CROWS * 64], o3 = cr[3 * AM_CROWS * 64], c2 = cr[4 * AM_CROWS * 64];
    float2 nx, nz;
    {
      const float d1 = pc.y - pc.x;
      const float d2 = pr0 - pl1;
      const float b = kC1 * d1 + kC2 * d2;
      const float u = o2.x * vx.x;
      nx.x = u - c2.x * b;
    }
    {
      const float d1 = pr0 - pc.y;
      const float d2 = pr1 - pc.x;
      const float b = kC1 * d1 + kC2 * d2;
      const float u = o2.y * vx.y;
      nx.y = u - c2.y * b;
    }
    {
      const float d1 = pd1.x - pc.x;
      const float d2 = pd2.x - pu.x;
      const float b = kC1 * d1 + kC2 * d2;
      const float u = o3.x * vz.x;
      nz.x = u - c2.x * b;
    }
    {
      const float d1 = pd1.y - pc.y;
      const float d2 = pd2.y - pu.y;
      const float b = kC1 * d1 + kC2 * d2;
      const float u = o3.y * vz.y;
      nz.y = u - c2.y * b;
    }
    bool rx = true, rz = true;
    if constexpr (decltype(G)::value) { rx = row_x(r); rz = row_z(r); }
    const bool mx0 = rx && (!CM || cx[0]), mx1 = rx && (!CM || cx[1]);
    const bool mz0 = rz && (!CM || cz[0]), mz1 = rz && (!CM || cz[1]);
    if constexpr (CM || decltype(G)::value) {
      if (!mx0) nx.x = vx.x;
      if (!mx1) nx.y = vx.y;
      if (!mz0) nz.x = vz.x;
      if (!mz1) nz.y = vz.y;
    }
    if constexpr (decltype(G)::value) {
      if (!row_in(r)) { nx = f2zero(); nz = f2zero(); }
    }
    vx = nx;
    vz = nz;
  };
  auto src_add = [&](float2& pn, float supd) {
    if (x0 == ssx && cin[0]) pn.x = pn.x + supd;
    if (x0 + 1 == ssx && cin[1]) pn.y = pn.y + supd;
  };
  auto sq = [](const float2& v) { return make_float2(v.x * v.x, v.y * v.y); };

  // four-row register windows, slot = row & 3
  float2 WVZ[4], WVX[4], WP1[4], WVX1[4], WVZ1[4], WP2[4], WEP[4], WEX[4], WEZ[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    WVZ[k] = WVX[k] = WP1[k] = WVX1[k] = WVZ1[k] = WP2[k] = f2zero();
    WEP[k] = WEX[k] = WEZ[k] = f2zero();
  }
  // this shot's rows, loaded AM_PF iterations ahead: vz of row i+6, p and vx
  // of row i+5 (rows clamped into what the segment needs)
  const int ja = max(0, zs - 6), jb = min(d.NZ, ze + 6) - 1;
  float2 pf_p[AM_PF], pf_vx[AM_PF], pf_vz[AM_PF];
  auto prefetch = [&](auto K, int i) {
    constexpr int k = decltype(K)::value;
    const int j6 = min(max(i + 6, ja), jb), j5 = min(max(i + 5, ja), jb);
    pf_vz[k] = row_ld2(VZ0, j6, d.PX, xo4);
    pf_p[k] = row_ld2(P0, j5, d.PX, xo4);
    pf_vx[k] = row_ld2(VX0, j5, d.PX, xo4);
  };

  auto iter = [&](auto U, auto G, int base) {
    constexpr int uu = decltype(U)::value;
    constexpr bool GG = decltype(G)::value;
    const int i = base + uu;
    constexpr int pk = uu % AM_PF;
    float2 vz_in = pf_vz[pk], p_in = pf_p[pk], vx_in = pf_vx[pk];
    prefetch(std::integral_constant<int, pk>{}, i + AM_PF);
    if constexpr (GG) {
      if (!row_in(i + 6)) vz_in = f2zero();
      if (!row_in(i + 5)) { p_in = f2zero(); vx_in = f2zero(); }
    }
    if constexpr (CM) {
      if (!lane_ok) { vz_in = f2zero(); p_in = f2zero(); vx_in = f2zero(); }
    }
    // --- row i+6: the input vz ---
    WVZ[(uu + 6) & 3] = vz_in;
    // --- row r1 = i+5: p' of step it ---
    {
      const int r1 = i + 5;
      float2 br;
      float2 pn = p_upd(G, r1, sCl + (r1 & (AM_CROWS - 1)) * 64, p_in, vx_in, WVZ[(uu + 3) & 3],
                        WVZ[(uu + 4) & 3], WVZ[(uu + 5) & 3], WVZ[(uu + 6) & 3], br);
      if constexpr (GG) {
        if (r1 == ssz) src_add(pn, supd0);
      }
      WP1[(uu + 5) & 3] = pn;
      WVX[(uu + 5) & 3] = vx_in;
      if constexpr (RECOMP) {
        if ((!GG || seg_row(r1)) && own_lane) row_st2(D0, r1, d.PX, xo4, br);
      }
    }
    // --- row r3 = i+3: vx', vz' of step it; its receivers and energy squares ---
    {
      constexpr int s3 = (uu + 3) & 3;
      const int r3 = i + 3;
      float2 vx = WVX[s3], vz = WVZ[s3];
      v_upd(G, r3, sCl + (r3 & (AM_CROWS - 1)) * 64, WP1[(uu + 2) & 3], WP1[s3], WP1[uu & 3],
            WP1[(uu + 1) & 3], vx, vz);
      WVX1[s3] = vx;
      WVZ1[s3] = vz;
      if constexpr (!RECOMP) {
        WEP[s3] = sq(WP1[s3]);
        WEX[s3] = sq(vx);
        WEZ[s3] = sq(vz);
        if constexpr (GG) {
          if (rec && own_lane && seg_row(r3) && a.rrow[r3]) record(r3, it, WP1[s3], vx, vz);
        }
      }
    }
    // --- row r2 = i+2: p'' of step it+1 ---
    {
      const int r2 = i + 2;
      float2 br;
      float2 pn = p_upd(G, r2, sCl + (r2 & (AM_CROWS - 1)) * 64, WP1[(uu + 2) & 3],
                        WVX1[(uu + 2) & 3], WVZ1[uu & 3], WVZ1[(uu + 1) & 3], WVZ1[(uu + 2) & 3],
                        WVZ1[(uu + 3) & 3], br);
      if constexpr (GG) {
        if (r2 == ssz) src_add(pn, supd1);
      }
      WP2[(uu + 2) & 3] = pn;
      if constexpr (RECOMP) {
        if ((!GG || seg_row(r2)) && own_lane) row_st2(D1, r2, d.PX, xo4, br);
      }
    }
    // --- row i: vx'', vz'' of step it+1; the row is stored ---
    {
      constexpr int s0 = uu & 3;
      float2 vx = WVX1[s0], vz = WVZ1[s0];
      const float2 pc = WP2[s0];
      v_upd(G, i, sCl + (i & (AM_CROWS - 1)) * 64, WP2[(uu + 3) & 3], pc, WP2[(uu + 1) & 3],
            WP2[(uu + 2) & 3], vx, vz);
      if (!GG || seg_row(i)) {
        if (own_lane) {
          row_st2(P1, i, d.PX, xo4, pc);
          row_st2(VX1, i, d.PX, xo4, vx);
          row_st2(VZ1, i, d.PX, xo4, vz);
        }
        if constexpr (!RECOMP) {
          if constexpr (GG) {
            if (rec && own_lane && a.rrow[i]) record(i, it + 1, pc, vx, vz);
          }
          const int sl = i & (FM_ESLOTS - 1);
          const float2 ep = sq(pc), ex = sq(vx), ez = sq(vz);
          sE[0][w][sl][lane] = make_float2(WEP[s0].x + ep.x, WEP[s0].y + ep.y);
          sE[1][w][sl][lane] = make_float2(WEX[s0].x + ex.x, WEX[s0].y + ex.y);
          sE[2][w][sl][lane] = make_float2(WEZ[s0].x + ez.x, WEZ[s0].y + ez.y);
        }
      }
    }
  };

  // coefficient ring: per block 5 planes x 4 rows, one row-plane per wavefront
  // at a time (uniform row base, lane offset)
  constexpr int CRP = 20, CPW = (CRP + NW - 1) / NW;
  float2 cpf[CPW];
  auto coef_load = [&](int r0) {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      const int rp = w + k * NW;  // plane * 4 + row
      if (rp < CRP) {
        const int r = min(max(r0 + (rp & 3), 0), d.NZ - 1);
        cpf[k] = row_ld2(a.coef + (rp >> 2) * d.plane, r, d.PX, xo4);
      }
    }
  };
  auto coef_store = [&](int r0) {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      const int rp = w + k * NW;
      if (rp < CRP) sC[rp >> 2][(r0 + (rp & 3)) & (AM_CROWS - 1)][lane] = cpf[k];
    }
  };
  const int i0 = zs - 12;
  // rows i0 .. i0+11 for the first block
  coef_load(i0); coef_store(i0);
  coef_load(i0 + 4); coef_store(i0 + 4);
  coef_load(i0 + 8); coef_store(i0 + 8);
  __syncthreads();
  if (active) {
    prefetch(std::integral_constant<int, 0>{}, i0);
    if constexpr (AM_PF > 1) prefetch(std::integral_constant<int, 1 % AM_PF>{}, i0 + 1);
    if constexpr (AM_PF > 2) {
      prefetch(std::integral_constant<int, 2 % AM_PF>{}, i0 + 2);
      prefetch(std::integral_constant<int, 3 % AM_PF>{}, i0 + 3);
    }
  }

  constexpr int RED = 4 * 128 / (NW * 64);  // reduced cells per thread per block and field
  const long long inner = (long long)d.nz * d.nx;
  float* __restrict__ eg = (!RECOMP && a.energy) ? a.energy + (long long)group * 3 * inner : nullptr;
  for (int base = i0; base < ze; base += 4) {
    // coefficient rows of the next block (base+12..base+15; this block uses
    // base..base+8, the ring rows base+9 .. base+11 are already there)
    coef_load(base + 12);
    // the energy partials this block completes (rows base..base+3), loaded early
    float ev[3][RED];
    long long eo[RED];
    bool eok[RED];
#pragma unroll
    for (int k = 0; k < RED; ++k) {
      const int q = threadIdx.x + k * NW * 64;
      const int row = base + (q >> 7), cl = q & 127, col = xs + cl;
      eok[k] = eg && row >= zs && row < ze && cl >= 6 && cl < 6 + AM_OWN && col < d.NX &&
               row >= d.nabc && row < d.nabc + d.nz && col >= d.nabc && col < d.nabc + d.nx;
      eo[k] = eok[k] ? (long long)(row - d.nabc) * d.nx + (col - d.nabc) : 0;
#pragma unroll
      for (int f = 0; f < 3; ++f) ev[f][k] = eok[k] ? eg[f * inner + eo[k]] : 0.f;
    }
    if (active) {
      // rows this block's stages touch: base .. base+9
      const int rr = base + lane;
      const bool rcv = rec && lane < 10 && rr >= 0 && rr < d.NZ && a.rrow[rr];
      const bool plain = base >= zs && base + 9 < ze && base >= d.f + 1 && base + 9 <= d.NZ - 3 &&
                         __ballot(rcv) == 0 && !(ssz >= base && ssz <= base + 9);
      if (plain) {
        iter(std::integral_constant<int, 0>{}, std::false_type{}, base);
        iter(std::integral_constant<int, 1>{}, std::false_type{}, base);
        iter(std::integral_constant<int, 2>{}, std::false_type{}, base);
        iter(std::integral_constant<int, 3>{}, std::false_type{}, base);
      } else {
        iter(std::integral_constant<int, 0>{}, std::true_type{}, base);
        iter(std::integral_constant<int, 1>{}, std::true_type{}, base);
        iter(std::integral_constant<int, 2>{}, std::true_type{}, base);
        iter(std::integral_constant<int, 3>{}, std::true_type{}, base);
      }
    }
    coef_store(base + 12);
    __syncthreads();
    if constexpr (!RECOMP) {
      // energy: each cell's squares summed over the group's shots in shot order
#pragma unroll
      for (int k = 0; k < RED; ++k) {
        if (!eok[k]) continue;
        const int q = threadIdx.x + k * NW * 64;
        const int row = base + (q >> 7), cl = q & 127;
        const int slot = row & (FM_ESLOTS - 1);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          const float* t = reinterpret_cast<const float*>(&sE[f][0][slot][0]) + cl;
          float acc = 0.f;
          for (int ww = 0; ww < ns; ++ww) acc = acc + t[ww * FM_ESLOTS * 128];
          eg[f * inner + eo[k]] = ev[f][k] + acc;
        }
      }
    }
  }
}

template <int NW, bool RECOMP>
__global__ __launch_bounds__(NW * 64) void fwd2m_kernel(FwdArgs a) {
  __shared__ float2 sE[RECOMP ? 1 : 3][NW][FM_ESLOTS][64];  // energy squares: field, shot, row slot, lane
  __shared__ float2 sC[5][AM_CROWS][64];                     // coefficient ring: plane, row slot, lane
  const int groups = (a.S + a.spb - 1) / a.spb;
  const int u = xcd_tile();
  if (u >= a.ntiles * groups) return;  // uniform per workgroup
  const int group = u % groups;
  const int seg = (u / groups) % a.am_segs;
  const int strip = u / groups / a.am_segs;
  const int xs = strip * AM_OWN - 6;
  if (xs >= 2 && xs + 127 <= a.d.NX - 3)
    fwd2m_body<NW, false, RECOMP>(a, group, seg, strip, sE, sC);
  else
    fwd2m_body<NW, true, RECOMP>(a, group, seg, strip, sE, sC);
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

// ---------------------------------------------------------------------------
// Batched gradient of uneven problems
//
// Problems are aligned at their last step: problem i's local step t runs at
// global step g = t + off_i, off_i = T - nt_i, T the longest step count
// rounded up to even.  Checkpoints sit at global multiples of k (even), so
// every problem's checkpoint intervals ("segments") coincide and one launch per
// two global steps covers every problem that has started: forward, segment
// recompute and adjoint keep the single-problem launch structure.  Sorted by
// step count, longest first, the problems running at a global step are a
// prefix, and their work units are numbered consecutively as in the batched
// forward.  A problem whose first step falls on the second step of a launch
// skips the first one (skip1; its state is still zero); in reverse, a launch
// whose second step would precede the problem's first runs one step (one).
// ---------------------------------------------------------------------------
// forward (RECOMP false, global steps g, g+1; launch j of the forward) or the
// recompute of segment c (RECOMP true, launch j of the segment, g = c*k + 2j)
template <int TZ, int TXT, bool RECOMP>
__global__ __launch_bounds__(FWD2_THREADS) void grad_batch_step_kernel(
    const GradDesc* __restrict__ descs, const int* __restrict__ unit_start, int n_active, int g,
    int j, int c, int k) {
  const int u = xcd_tile();
  if (u >= unit_start[n_active]) return;
  const int p = batch_problem(unit_start, n_active, u);
  const GradDesc& bd = descs[p];
  FwdArgs a = bd.f;
  const float* in;
  float* out;
  if (RECOMP) {
    out = (j & 1) ? bd.B : bd.A;
    in = j == 0 ? reinterpret_cast<const float*>(bd.slots + (long long)(c - bd.c0) * bd.slot_bytes)
                : ((j & 1) ? bd.A : bd.B);
    a.dbuf = bd.D + (long long)(g - c * k) * bd.fstride;
    a.bstride = bd.fstride;
    a.gxbuf = a.gzbuf = nullptr;
  } else {
    in = (j & 1) ? bd.B : bd.A;
    out = (j & 1) ? bd.A : bd.B;
  }
  a.P0 = in; a.VX0 = in + bd.fstride; a.VZ0 = in + 2 * bd.fstride;
  a.P1 = out; a.VX1 = out + bd.fstride; a.VZ1 = out + 2 * bd.fstride;
  a.it = g - bd.off;
  a.skip1 = a.it == -1;
  a.one = 0;
  const int local = u - unit_start[p];
  fwd2_group<TZ, TXT, FWD2_THREADS, RECOMP, true>(a, local % a.ntiles, local / a.ntiles);
}

// per-problem fills and copies of the batched gradient (blockIdx.y = problem):
//   0: forward start: zero A, B and the first checkpoint slot
//   1: segment c start: zero A and B of the problems whose first step lies in c
//   2: reverse start: zero the adjoint state (both sides), imaging and wavelet sums
//   3: checkpoint c: copy the state in A (which = 0) or B (which = 1) to slot c
__global__ void grad_batch_prep_kernel(const GradDesc* __restrict__ descs, int mode, int c,
                                       int which) {
  const GradDesc& bd = descs[blockIdx.y];
  const long long st = 3 * bd.fstride;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 0) {
    float* s0 = reinterpret_cast<float*>(bd.slots);
    for (long long i = i0; i < st; i += stride) { bd.A[i] = 0.f; bd.B[i] = 0.f; s0[i] = 0.f; }
  } else if (mode == 1) {
    if (bd.c0 != c) return;
    for (long long i = i0; i < st; i += stride) { bd.A[i] = 0.f; bd.B[i] = 0.f; }
  } else if (mode == 2) {
    for (long long i = i0; i < st; i += stride) { bd.LA[i] = 0.f; bd.LB[i] = 0.f; }
    const long long ng = (long long)bd.groups * bd.b.d.plane;
    for (long long i = i0; i < ng; i += stride) bd.G1[i] = 0.f;
    const long long nw = (long long)bd.b.S * bd.b.nt;
    for (long long i = i0; i < nw; i += stride) bd.GW[i] = 0.f;
  } else {
    if (c <= bd.c0) return;  // slot c0 holds the zero state from the start
    const float* src = which ? bd.B : bd.A;
    float* dst = reinterpret_cast<float*>(bd.slots + (long long)(c - bd.c0) * bd.slot_bytes);
    for (long long i = i0; i < st; i += stride) dst[i] = src[i];
  }
}

// lambda at each problem's last step: scatter of that step's adjoint sources
// into LA (zeroed), as adj_inject_kernel does for one problem
__global__ void grad_batch_inject_kernel(const GradDesc* __restrict__ descs) {
  const GradDesc& bd = descs[blockIdx.y];
  const AdjArgs& b = bd.b;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= b.S * b.R) return;
  const int s = idx / b.R, r = idx % b.R;
  const int cell = bd.rcell[r];
  if (b.rhead[cell] != r) return;  // only the list head sums its cell
  float sp = 0.f, sx = 0.f, sz = 0.f;
  for (int h = r; h >= 0; h = b.rnext[h]) {
    const long long o = ((long long)s * b.nt + (b.nt - 1)) * b.R + h;
    if (b.g_p) sp += b.g_p[o];
    if (b.g_vx) sx += b.g_vx[o];
    if (b.g_vz) sz += b.g_vz[o];
  }
  const long long q = (long long)s * b.d.plane + cell;
  bd.LA[q] = bd.LA[q] + sp;
  bd.LA[bd.fstride + q] = bd.LA[bd.fstride + q] + sx;
  bd.LA[2 * bd.fstride + q] = bd.LA[2 * bd.fstride + q] + sz;
}

__global__ void stamp_kernel(unsigned long long* t) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *t = wall_clock64();
}

// chain rule alpha1 = den*c*c*dt/dz, alpha2 = dt/(den*dz) and the fold of
// pad_model's edge replication (acoustic_2d.py:407-410, :18-66)
__device__ __forceinline__ void fold_cell(const Dims& d, int groups, const float* __restrict__ gA1,
                                          const float* __restrict__ gA2,
                                          const float* __restrict__ vp,
                                          const float* __restrict__ rho, float dt, float dz,
                                          float* __restrict__ grad_vp,
                                          float* __restrict__ grad_rho, int a, int b) {
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

__global__ void fold_kernel(Dims d, int groups, const float* __restrict__ gA1,
                            const float* __restrict__ gA2, const float* __restrict__ vp,
                            const float* __restrict__ rho, float dt, float dz,
                            float* __restrict__ grad_vp, float* __restrict__ grad_rho) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= d.nx || a >= d.nz) return;
  fold_cell(d, groups, gA1, gA2, vp, rho, dt, dz, grad_vp, grad_rho, a, b);
}

// wavelet gradient: per shot, or summed over shots for a shared wavelet
__device__ __forceinline__ void wavelet_grad(const float* __restrict__ gw, int S, int nt,
                                             int shared, float* __restrict__ out, int t) {
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

__global__ void wavelet_grad_kernel(const float* __restrict__ gw, int S, int nt, int shared,
                                    float* __restrict__ out) {
  wavelet_grad(gw, S, nt, shared, out, blockIdx.x * blockDim.x + threadIdx.x);
}

// several points per shot, one wavelet per shot: the samples of its K points
// (rows s*K .. s*K+K-1 of gw) summed in point order
__global__ void wavelet_grad_points_kernel(const float* __restrict__ gw, int S, int K, int nt,
                                           float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)S * nt) return;
  const long long s = t / nt, i = t - s * nt;
  float v = gw[s * K * nt + i];
  for (int k = 1; k < K; ++k) v = v + gw[(s * K + k) * nt + i];
  out[t] = v;
}

// a problem of the batched gradient is done: chain rule and pad_model fold of
// its imaging sums (blockIdx.y = model row), then its wavelet gradient
__global__ void grad_batch_fold_kernel(const GradDesc* __restrict__ descs, int p) {
  const GradDesc& bd = descs[p];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  const Dims& d = bd.f.d;
  if (!bd.grad_vp || b >= d.nx || a >= d.nz) return;
  fold_cell(d, bd.groups, bd.G1, nullptr, bd.vp, bd.rho, bd.f.dt, bd.dz, bd.grad_vp, nullptr, a, b);
}

__global__ void grad_batch_wgrad_kernel(const GradDesc* __restrict__ descs, int p) {
  const GradDesc& bd = descs[p];
  if (!bd.grad_w) return;
  wavelet_grad(bd.GW, bd.b.S, bd.b.nt, bd.shared, bd.grad_w, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// BF16 checkpoints: per (field, shot) plane, a power-of-two FP32 scale s >=
// max|x| and bf16(x / s) rounded to nearest even on the FP32 bit pattern.
// Dividing and multiplying by a power of two is exact, so the only rounding
// is the bf16 one (8 significant bits).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// power-of-two scale >= m (m >= 0): 2^(E-126) for biased exponent E of m
__device__ __forceinline__ float pow2_scale(float m) {
  const unsigned e = (__float_as_uint(m) >> 23) & 0xffu;
  if (!(m > 0.f) || e >= 254u) return 1.f;
  return __uint_as_float((e + 1u) << 23);
}

// maxbits[plane] = max |x| over each plane (bits of a non-negative float order
// like unsigned integers, so an integer max is a float max)
__global__ void absmax_kernel(long long plane, const float* __restrict__ st,
                              unsigned* __restrict__ maxbits) {
  __shared__ float red[NTHREADS / 64];
  const float* p = st + (long long)blockIdx.y * plane;
  float m = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
       i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(p[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NTHREADS / 64; ++w) m = fmaxf(m, red[w]);
    atomicMax(&maxbits[blockIdx.y], __float_as_uint(m));
  }
}

__global__ void quantize_kernel(long long plane, const float* __restrict__ st,
                                const unsigned* __restrict__ maxbits,
                                unsigned short* __restrict__ q, float* __restrict__ scales) {
  const float scale = pow2_scale(__uint_as_float(maxbits[blockIdx.y]));
  const float inv = 1.f / scale;
  const float* p = st + (long long)blockIdx.y * plane;
  unsigned short* o = q + (long long)blockIdx.y * plane;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
       i += (long long)gridDim.x * blockDim.x)
    o[i] = f32_to_bf16(p[i] * inv);
  if (blockIdx.x == 0 && threadIdx.x == 0) scales[blockIdx.y] = scale;
}

__global__ void dequantize_kernel(long long plane, const unsigned short* __restrict__ q,
                                  const float* __restrict__ scales, float* __restrict__ st) {
  const float scale = scales[blockIdx.y];
  const unsigned short* p = q + (long long)blockIdx.y * plane;
  float* o = st + (long long)blockIdx.y * plane;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
       i += (long long)gridDim.x * blockDim.x)
    o[i] = bf16_to_f32(p[i]) * scale;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Plan {
  Dims d;
  int S, R, nt, k, nseg, TZ, spb, groups, tiles_x, tiles_z, echunk;
  int K;            // source points per shot
  int rspb, rgroups;  // shot groups of the gradient's tiled recompute (the adjoint's may be smaller)
  bool bf16;
  bool march;       // z-marching recompute instead of tiles
  bool amarch;      // z-marching two-step adjoint (adj2m_kernel) instead of tiles
  bool fmarch;      // z-marching two-step forward / recompute (fwd2m_kernel) instead of tiles
  bool mixed;       // two-phase two-step forward / recompute launches (fwd2_mixed_kernel)
  int mix_ta;       // their phase-A tiles
  int eparts;       // energy partial maps (shot groups; the two-phase launches' larger count)
  int am_strips, am_segs, am_seg_h;  // its column strips, z segments and rows per segment
  int am_seg0_h;    // the adjoint's first segment (shorter with a free surface)
  bool two;         // two time steps per forward / recompute launch
  int TZ2, TX2, tiles_x2;  // tiles of the two-step kernels (16 x 128 or 24 x 96)
  int mseg, mseg_h; // z segments per strip and rows per segment for it
  size_t plane_bytes, slot_bytes, slot_scale_off;
  // byte offsets into the workspace
  size_t off_coef, off_A, off_B, off_LA, off_LB, off_D, off_GX, off_GZ, off_E, off_G1, off_G2,
      off_GW, off_geo, off_MX, total;
  // geometry region (receiver maps, source cells): offsets from its start; it
  // lies in the workspace at off_geo, or in a caller's geometry buffer
  // (gbw_acoustic2d_geometry_setup)
  size_t off_rhead, off_rnext, off_rrow, off_rcell, off_heads, off_src, geo_total;
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
  Dims& d = pl.d;
  d.nz = p->nz; d.nx = p->nx; d.nabc = p->nabc;
  d.NZ = p->nz + 2 * p->nabc; d.NX = p->nx + 2 * p->nabc;
  d.PX = (d.NX + 63) / 64 * 64;
  d.fs = p->free_surface ? 1 : 0;
  d.f = d.fs ? p->nabc : 1;
  d.plane = (long long)d.NZ * d.PX;
  if (d.NZ < 6 || d.NX < 6) return fail(GBW_ERR_INVALID, "padded grid must be at least 6 x 6");
  // in-plane float4 accesses use 32-bit byte offsets
  if (d.plane >= (1LL << 30)) return fail(GBW_ERR_INVALID, "padded plane exceeds 2^30 cells");
  pl.S = p->src_n; pl.R = p->rcv_n; pl.nt = p->nt;
  if (p->src_points < 0) return fail(GBW_ERR_INVALID, "src_points must be >= 0");
  pl.K = std::max(1, p->src_points);
  if (p->wavelet_shared < 0 || p->wavelet_shared > 2)
    return fail(GBW_ERR_INVALID, "wavelet_shared must be 0, 1 or 2");
  if (pl.K > 1 && p->wavelet_shared == 1)
    return fail(GBW_ERR_INVALID,
                "several source points per shot take src_v of shape (src_n, nt) or "
                "(src_n, src_points, nt)");
  pl.k = p->checkpoint_every > 0 ? std::min(p->checkpoint_every, p->nt) : 0;
  if (pass >= 1 && pl.k == 0) return fail(GBW_ERR_INVALID, "the gradient needs checkpoint_every > 0");
  pl.nseg = pl.k ? (p->nt + pl.k - 1) / pl.k : 0;
  pl.TZ = p->tile_rows == 32 ? 32 : 16;
  if (p->tile_rows != 0 && p->tile_rows != 16 && p->tile_rows != 24 && p->tile_rows != 32)
    return fail(GBW_ERR_INVALID, "tile_rows must be 0, 16, 24 or 32");
  // two-step kernels: 24 rows x 96 columns on request, else 16 x 128
  pl.TZ2 = p->tile_rows == 24 ? 24 : 16;
  pl.TX2 = pl.TZ2 == 24 ? 96 : 128;
  pl.tiles_x2 = (d.NX + pl.TX2 - 1) / pl.TX2;
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
  if (p->steps_per_launch != 1) spb = std::min(spb, MAX_SPB);
  pl.spb = spb;
  pl.groups = (pl.S + spb - 1) / spb;
  pl.echunk = p->energy_chunk > 0 ? std::min(p->energy_chunk, p->nt) : p->nt;
  if (p->march < 0 || p->march > 3) return fail(GBW_ERR_INVALID, "march must be 0, 1, 2 or 3");
  if (p->steps_per_launch < 0 || p->steps_per_launch > 2)
    return fail(GBW_ERR_INVALID, "steps_per_launch must be 0, 1 or 2");
  // several source points per shot: one-step tiled launches (the two-step and
  // z-marching kernels inject one point per shot)
  pl.two = p->steps_per_launch != 1 && pl.K == 1;
  // z-marching recompute only on request: with coefficients reloaded per shot
  // it measured slower than tiles on 128-column strips
  pl.march = p->march == 2 && pl.K == 1;
  // z-marching two-step adjoint: one wavefront per shot of a group of at most 8;
  // chosen by default for gradients of 8 or more shots without the density
  // gradient (it fills its workgroups and measured faster than tiles: 64 shots
  // of 390 x 1740, 426 vs 622 us per launch)
  pl.amarch = pl.two && pass >= 1 &&
              (p->march == 3 || (p->march == 0 && pass == 1 && pl.S >= 8));
  if (p->fwd_march < 0 || p->fwd_march > 3)
    return fail(GBW_ERR_INVALID, "fwd_march must be 0, 1, 2 or 3");
  // z-marching two-step forward and recompute: one wavefront per shot of a
  // group of at most 8, on request; without a free surface, and not for the
  // density gradient's recompute (its extra brackets keep the tiles)
  pl.fmarch = pl.two && !d.fs && pass != 2 && p->fwd_march == 2;
  // the tiled recompute keeps the requested shot groups when only the adjoint
  // marches (its workgroups take at most 8 shots)
  pl.rspb = pl.spb;
  if (pl.amarch || pl.fmarch) {
    pl.spb = std::min(pl.spb, 8);
    pl.groups = (pl.S + pl.spb - 1) / pl.spb;
  }
  if (pl.fmarch) pl.rspb = pl.spb;
  pl.rgroups = (pl.S + pl.rspb - 1) / pl.rspb;
  // two-phase two-step launches: phase A takes the largest multiple of the CU
  // count of tiles, phase B the rest in groups of MIX_SPB_B shots; on request
  // (fwd_march 3), and by default where phase A's groups are the larger
  // (measured 334 -> 316 us per forward and 348 -> 326 us per recompute launch,
  // 64 shots of 390 x 1740)
  {
    const int t2 = pl.tiles_x2 * ((d.NZ + pl.TZ2 - 1) / pl.TZ2);
    pl.mix_ta = t2 / AM_TARGET_WGS * AM_TARGET_WGS;
    const bool fits = pl.two && !pl.fmarch && pl.mix_ta > 0 && pl.mix_ta < t2;
    const int spb = pass == 0 ? pl.spb : pl.rspb;  // the two-step launches' groups
    pl.mixed = fits && (p->fwd_march == 3 || (p->fwd_march == 0 && spb > MIX_SPB_B));
    pl.eparts = pl.groups;
    if (pl.mixed && pass == 0)
      pl.eparts = std::max(pl.groups, (pl.S + mix_spb_b(pl.spb) - 1) / mix_spb_b(pl.spb));
  }
  pl.am_strips = (d.NX + AM_OWN - 1) / AM_OWN;
  {
    // z segments of the marching kernels: one workgroup runs per CU (its
    // registers and LDS allow no second), so a launch takes about
    // ceil(workgroups / CUs) rounds of one segment each, and a segment costs
    // its rows plus the 12-row pipeline start; take the count that minimises
    // that (segments of at least 16 rows)
    const int per_seg = pl.am_strips * pl.groups;
    int best_ns = 1;
    long long best = -1;
    for (int ns = 1; ns <= std::max(1, d.NZ / 16); ++ns) {
      const int h = ((d.NZ + ns - 1) / ns + 3) / 4 * 4;
      const long long units = (long long)per_seg * ((d.NZ + h - 1) / h);
      const long long cost = (units + AM_TARGET_WGS - 1) / AM_TARGET_WGS * (h + 12);
      if (best < 0 || cost < best) { best = cost; best_ns = ns; }
    }
    pl.am_seg_h = ((d.NZ + best_ns - 1) / best_ns + 3) / 4 * 4;
    pl.am_segs = (d.NZ + pl.am_seg_h - 1) / pl.am_seg_h;
    pl.am_seg0_h = pl.am_seg_h;
    // The adjoint's first segment holds the rows a survey's sources and
    // receivers usually sit on, and with a free surface the rows above the
    // surface: blocks there run in the guarded (slower) form.  In a launch of
    // one round, where the longest segment sets the time, it takes `shift`
    // rows fewer than the others (64 shots of 390 x 1740, receivers and
    // sources on row 2: 8 rows fewer 425 -> 416 us per launch, 4 and 16 fewer
    // 421 and 422 us; with a free surface f = 20 rows fewer 471 -> 455 us,
    // f/2 and 3f/2 fewer 459 and 461 us).  The segment count stays.
    const int shift = d.fs ? std::max(d.f, 8) : 8;
    if (pl.am_segs >= 2 && (long long)per_seg * pl.am_segs <= AM_TARGET_WGS &&
        4 * shift <= pl.am_seg_h) {
      const int ns = pl.am_segs;
      const int h0 = std::max(16, (d.NZ - (ns - 1) * shift) / ns / 4 * 4);
      const int h = ((d.NZ - h0 + ns - 2) / (ns - 1) + 3) / 4 * 4;
      if (h0 < h && h0 + (ns - 1) * h >= d.NZ) {
        pl.am_seg0_h = h0;
        pl.am_seg_h = h;
        pl.am_segs = 1 + (d.NZ - h0 + h - 1) / h;
      }
    }
  }
  {
    const int strips = pl.tiles_x;
    int ns = std::max(1, std::min((2048 + strips * pl.S - 1) / (strips * pl.S), (d.NZ + 47) / 48));
    pl.mseg_h = (d.NZ + ns - 1) / ns;
    pl.mseg = (d.NZ + pl.mseg_h - 1) / pl.mseg_h;
  }
  pl.plane_bytes = (size_t)d.plane * sizeof(float);
  const size_t state = 3 * (size_t)pl.S * pl.plane_bytes;
  pl.bf16 = p->checkpoint_bf16 != 0;
  if (pl.bf16) {  // bf16 planes, then one fp32 scale per (field, shot)
    pl.slot_scale_off = align256(3 * (size_t)pl.S * d.plane * sizeof(unsigned short));
    pl.slot_bytes = pl.slot_scale_off + align256(3 * (size_t)pl.S * sizeof(float));
  } else {
    pl.slot_scale_off = 0;
    pl.slot_bytes = state;
  }
  const size_t inner = (size_t)d.nz * d.nx * sizeof(float);
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += align256(bytes); return at; };
  pl.off_coef = take(5 * pl.plane_bytes);
  pl.off_A = take(state);
  pl.off_B = take(state);
  pl.off_LA = pl.off_LB = pl.off_D = pl.off_GX = pl.off_GZ = pl.off_G1 = pl.off_G2 = pl.off_GW = 0;
  pl.off_E = 0;
  if (pass == 0) {
    pl.off_E = take((size_t)pl.eparts * 3 * inner);
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
    pl.off_GW = take((size_t)pl.S * pl.K * pl.nt * sizeof(float));
  }
  {
    size_t g = 0;
    auto gtake = [&](size_t bytes) { size_t at = g; g += align256(bytes); return at; };
    pl.off_rhead = gtake((size_t)d.plane * sizeof(int));
    pl.off_rnext = gtake((size_t)std::max(1, pl.R) * sizeof(int));
    pl.off_rcell = gtake((size_t)std::max(1, pl.R) * sizeof(int));
    pl.off_heads = gtake(2 * (size_t)std::max(1, pl.R) * sizeof(int));
    pl.off_rrow = gtake((size_t)d.NZ);
    pl.off_src = gtake((size_t)pl.S * pl.K * sizeof(int));
    pl.geo_total = g;
  }
  pl.off_geo = take(pl.geo_total);
  pl.off_MX = take(3 * (size_t)pl.S * sizeof(unsigned));
  pl.total = o;
  return 0;
}

size_t checkpoint_bytes(const Plan& pl) { return pl.k ? (size_t)pl.nseg * pl.slot_bytes : 0; }

unsigned plane_blocks(long long plane) { return (unsigned)std::min<long long>((plane + 255) / 256, 512); }

// state (fp32, [field][shot][plane]) -> bf16 checkpoint slot with scales
int quantize_state(const Plan& pl, const float* state, char* slot, unsigned* maxbits,
                   hipStream_t st) {
  const int planes = 3 * pl.S;
  fill_int_kernel<<<1, 256, 0, st>>>(reinterpret_cast<int*>(maxbits), planes, 0);
  LAUNCH_CHECK();
  dim3 g(plane_blocks(pl.d.plane), planes);
  absmax_kernel<<<g, NTHREADS, 0, st>>>(pl.d.plane, state, maxbits);
  LAUNCH_CHECK();
  quantize_kernel<<<g, 256, 0, st>>>(pl.d.plane, state, maxbits,
                                     reinterpret_cast<unsigned short*>(slot),
                                     reinterpret_cast<float*>(slot + pl.slot_scale_off));
  LAUNCH_CHECK();
  return 0;
}

int dequantize_state(const Plan& pl, const char* slot, float* state, hipStream_t st) {
  dim3 g(plane_blocks(pl.d.plane), 3 * pl.S);
  dequantize_kernel<<<g, 256, 0, st>>>(pl.d.plane, reinterpret_cast<const unsigned short*>(slot),
                                       reinterpret_cast<const float*>(slot + pl.slot_scale_off),
                                       state);
  LAUNCH_CHECK();
  return 0;
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
  // ws: the start of the geometry region
  const Dims& d = pl.d;
  if (!in || !in->src_x || !in->src_z) return fail(GBW_ERR_INVALID, "source positions are NULL");
  if (pl.R > 0 && (!in->rcv_x || !in->rcv_z)) return fail(GBW_ERR_INVALID, "receiver positions are NULL");
  // point k of shot s at s * K + k
  std::vector<int> src((size_t)pl.S * pl.K);
  for (int i = 0; i < pl.S * pl.K; ++i) {
    const long long z = in->src_z[i] + d.nabc, x = in->src_x[i] + d.nabc;
    const std::string who = pl.K > 1 ? "source " + std::to_string(i / pl.K) + " point " +
                                           std::to_string(i % pl.K)
                                     : "source " + std::to_string(i);
    if (z < 0 || z >= d.NZ || x < 0 || x >= d.NX)
      return fail(GBW_ERR_INDEX, "index out of range: " + who + " at (z=" +
                                     std::to_string(in->src_z[i]) + ", x=" +
                                     std::to_string(in->src_x[i]) + ")");
    src[i] = (int)(z * d.PX + x);
    // the reference's indexed update keeps one of two equal cells' sums
    // (acoustic_2d.py:237): such shots are refused, not guessed at
    for (int j = i - i % pl.K; j < i; ++j)
      if (src[j] == src[i])
        return fail(GBW_ERR_INVALID, who + " repeats the cell of point " +
                                         std::to_string(j % pl.K) + " of its shot");
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

// grid.x is the tile count rounded up to a multiple of 8 (the XCD count), so
// xcd_tile() is a bijection; surplus workgroups exit at once
inline unsigned tile_grid(const Plan& pl) { return (unsigned)((pl.tiles_x * pl.tiles_z + 7) / 8 * 8); }

// one forward / recompute launch of `steps` (1 or 2) time steps from a.it
// (the two-step kernel always tiles 16 rows, whatever tile_rows the one-step
// and adjoint kernels use)
template <bool RECOMP>
void launch_steps(const Plan& pl, const FwdArgs& a_in, int steps, hipStream_t st) {
  // shot groups: the recompute's own (rspb), else the pass's
  FwdArgs a = a_in;
  a.spb = RECOMP ? pl.rspb : pl.spb;
  const int groups = RECOMP ? pl.rgroups : pl.groups;
  dim3 grid(tile_grid(pl), groups);
  if (steps == 2 && pl.fmarch && !a.gxbuf && !a.one) {
    FwdArgs am = a;
    am.ntiles = pl.am_strips * pl.am_segs;
    am.am_segs = pl.am_segs;
    am.seg_h = pl.am_seg_h;
    const unsigned n = (unsigned)(((long long)am.ntiles * groups + 7) / 8 * 8);
    LaunchTimer timer(RECOMP ? "fwd2m_kernel<recompute>" : "fwd2m_kernel<forward>", st);
    if (pl.spb <= 4)
      fwd2m_kernel<4, RECOMP><<<n, 256, 0, st>>>(am);
    else
      fwd2m_kernel<8, RECOMP><<<n, 512, 0, st>>>(am);
    return;
  }
  if (steps == 2 && pl.mixed) {
    FwdArgs a2 = a;
    a2.tiles_x = pl.tiles_x2;
    a2.ntiles = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
    a2.mix_ta = pl.mix_ta;
    a2.mix_spb_b = mix_spb_b(a.spb);
    const int ga = (a.S + a.spb - 1) / a.spb, gb = (a.S + a2.mix_spb_b - 1) / a2.mix_spb_b;
    const int tb = a2.ntiles - pl.mix_ta;
    const unsigned n = 8u * (unsigned)(pl.mix_ta / 8 * ga + (tb * gb + 7) / 8);
    LaunchTimer timer(RECOMP ? "fwd2_mixed_kernel<recompute>" : "fwd2_mixed_kernel<forward>", st);
    if (pl.TZ2 == 24)
      fwd2_mixed_kernel<24, 96, FWD2_THREADS, RECOMP><<<n, FWD2_THREADS, 0, st>>>(a2);
    else
      fwd2_mixed_kernel<16, 128, FWD2_THREADS, RECOMP><<<n, FWD2_THREADS, 0, st>>>(a2);
    return;
  }
  LaunchTimer timer(steps == 2 ? (RECOMP ? "fwd2_step_kernel<recompute>" : "fwd2_step_kernel<forward>")
                               : (RECOMP ? "fwd_step_kernel<recompute>" : "fwd_step_kernel<forward>"),
                    st);
  if (steps == 2) {
    FwdArgs a2 = a;
    a2.tiles_x = pl.tiles_x2;
    a2.ntiles = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
    const dim3 g2((unsigned)((a2.ntiles + 7) / 8 * 8), groups);
    if (pl.TZ2 == 24)
      fwd2_step_kernel<24, 96, FWD2_THREADS, RECOMP><<<g2, FWD2_THREADS, 0, st>>>(a2);
    else
      fwd2_step_kernel<16, 128, FWD2_THREADS, RECOMP><<<g2, FWD2_THREADS, 0, st>>>(a2);
  } else if (pl.TZ == 32) {
    if (pl.K > 1)
      fwd_step_kernel<32, RECOMP, true><<<grid, NTHREADS, 0, st>>>(a);
    else
      fwd_step_kernel<32, RECOMP><<<grid, NTHREADS, 0, st>>>(a);
  } else {
    if (pl.K > 1)
      fwd_step_kernel<16, RECOMP, true><<<grid, NTHREADS, 0, st>>>(a);
    else
      fwd_step_kernel<16, RECOMP><<<grid, NTHREADS, 0, st>>>(a);
  }
}

// two adjoint steps (a.it, a.it-1) in one launch; 16-row tiles, or z-marching
// wavefronts (no density gradient)
void launch_adj2(const Plan& pl, const AdjArgs& a, hipStream_t st) {
  if (pl.amarch && !a.gA2) {
    AdjArgs am = a;
    am.ntiles = pl.am_strips * pl.am_segs;
    am.am_segs = pl.am_segs;
    am.am_seg_h = pl.am_seg_h;
    am.am_seg0_h = pl.am_seg0_h;
    const unsigned n = (unsigned)(((long long)am.ntiles * pl.groups + 7) / 8 * 8);
    LaunchTimer timer("adj2m_kernel", st);
    if (pl.d.fs) {
      if (pl.spb <= 4)
        adj2m_kernel<4, true><<<n, 256, 0, st>>>(am);
      else
        adj2m_kernel<8, true><<<n, 512, 0, st>>>(am);
    } else {
      if (pl.spb <= 4)
        adj2m_kernel<4, false><<<n, 256, 0, st>>>(am);
      else
        adj2m_kernel<8, false><<<n, 512, 0, st>>>(am);
    }
    return;
  }
  AdjArgs a2 = a;
  a2.tiles_x = pl.tiles_x2;
  a2.ntiles = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
  dim3 grid((unsigned)((a2.ntiles + 7) / 8 * 8), pl.groups);
  LaunchTimer timer("adj2_step_kernel", st);
  if (pl.TZ2 == 24) {
    if (a.gA2)
      adj2_step_kernel<24, 96, FWD2_THREADS, true><<<grid, FWD2_THREADS, 0, st>>>(a2);
    else
      adj2_step_kernel<24, 96, FWD2_THREADS, false><<<grid, FWD2_THREADS, 0, st>>>(a2);
  } else {
    if (a.gA2)
      adj2_step_kernel<16, 128, FWD2_THREADS, true><<<grid, FWD2_THREADS, 0, st>>>(a2);
    else
      adj2_step_kernel<16, 128, FWD2_THREADS, false><<<grid, FWD2_THREADS, 0, st>>>(a2);
  }
}

template <int TZ>
void launch_adj(const Plan& pl, const AdjArgs& a, hipStream_t st) {
  dim3 grid(tile_grid(pl), pl.groups);
  LaunchTimer timer("adj_step_kernel", st);
  if (pl.K > 1) {
    if (a.gA2)
      adj_step_kernel<TZ, true, true><<<grid, NTHREADS, 0, st>>>(a);
    else
      adj_step_kernel<TZ, false, true><<<grid, NTHREADS, 0, st>>>(a);
  } else if (a.gA2) {
    adj_step_kernel<TZ, true><<<grid, NTHREADS, 0, st>>>(a);
  } else {
    adj_step_kernel<TZ, false><<<grid, NTHREADS, 0, st>>>(a);
  }
}

// Prepared geometry buffers (gbw_acoustic2d_geometry_setup), by device
// address, with the shape each was prepared for: a call handed one checks its
// own shape against it before any kernel reads the buffer.
struct GeoKey {
  int NZ, NX, PX, nabc, fs, S, R, K;
  bool operator==(const GeoKey& o) const {
    return NZ == o.NZ && NX == o.NX && PX == o.PX && nabc == o.nabc && fs == o.fs && S == o.S &&
           R == o.R && K == o.K;
  }
};
std::mutex g_geo_mu;
std::unordered_map<const void*, GeoKey> g_geo;

GeoKey geo_key(const Plan& pl) {
  return GeoKey{pl.d.NZ, pl.d.NX, pl.d.PX, pl.d.nabc, pl.d.fs, pl.S, pl.R, pl.K};
}

// the geometry region of a call: the caller's prepared buffer, or the
// workspace's, built now from the positions (which synchronises `st`)
int resolve_geometry(const Plan& pl, const gbw_acoustic2d_inputs* in, void* ws, hipStream_t st,
                     const void** geo) {
  if (!in) return fail(GBW_ERR_INVALID, "inputs is NULL");
  if (in->geometry) {
    std::lock_guard<std::mutex> lk(g_geo_mu);
    auto it = g_geo.find(in->geometry);
    if (it == g_geo.end() || !(it->second == geo_key(pl)))
      return fail(GBW_ERR_INVALID,
                  "geometry buffer was not prepared by gbw_acoustic2d_geometry_setup for this "
                  "problem's grid, shots and receivers");
    *geo = in->geometry;
    return 0;
  }
  void* g = static_cast<char*>(ws) + pl.off_geo;
  std::vector<int> rcell;
  int rc = setup_geometry(pl, in, g, st, rcell);
  if (rc) return rc;
  *geo = g;
  return 0;
}

// geo: the geometry region (in the workspace or a caller's geometry buffer)
FwdArgs base_fwd_args(const Plan& pl, const gbw_acoustic2d_problem* p,
                      const gbw_acoustic2d_inputs* in, void* ws, const void* geo) {
  void* g = const_cast<void*>(geo);
  FwdArgs a{};
  a.d = pl.d; a.S = pl.S; a.spb = pl.spb; a.tiles_x = pl.tiles_x;
  a.ntiles = pl.tiles_x * pl.tiles_z;
  a.coef = at<float>(ws, pl.off_coef);
  a.src_cell = at<int>(g, pl.off_src);
  a.wav = in->src_v;
  // wavelet rows: shared (nt,), per shot (src_n, nt), per point (src_n, K, nt)
  a.wav_stride = p->wavelet_shared == 1 ? 0 : (p->wavelet_shared == 2 ? (long long)pl.K : 1) * p->nt;
  a.wav_kstride = p->wavelet_shared == 2 ? p->nt : 0;
  a.src_k = pl.K;
  a.nt = p->nt; a.dt = p->dt;
  a.rhead = at<int>(g, pl.off_rhead);
  a.rnext = at<int>(g, pl.off_rnext);
  a.rrow = at<unsigned char>(g, pl.off_rrow);
  a.R = pl.R;
  return a;
}

void set_state(FwdArgs& a, float* in, float* out, long long fstride) {
  a.P0 = in; a.VX0 = in + fstride; a.VZ0 = in + 2 * fstride;
  a.P1 = out; a.VX1 = out + fstride; a.VZ1 = out + 2 * fstride;
}

// ---------------------------------------------------------------------------
// Batched forward: layout and schedule (host)
// ---------------------------------------------------------------------------
struct BatchLayout {
  int n = 0, max_nt = 0, single = 0;
  std::vector<int> order;        // sorted position -> caller's problem index
  std::vector<Plan> plans;       // per sorted position
  std::vector<size_t> off_prob;  // per sorted position: its sub-workspace
  std::vector<int> unit_start;   // n + 1 prefix sums of (two-step tiles x shot groups)
  size_t off_desc = 0, off_units = 0, off_counters = 0, total = 0;
};

int make_batch(const gbw_acoustic2d_problem* probs, int n, BatchLayout& L) {
  if (!probs || n <= 0) return fail(GBW_ERR_INVALID, "batch needs n > 0 problems");
  L = BatchLayout();
  L.n = n;
  L.order.resize(n);
  for (int i = 0; i < n; ++i) L.order[i] = i;
  std::stable_sort(L.order.begin(), L.order.end(),
                   [&](int x, int y) { return probs[x].nt > probs[y].nt; });
  L.single = probs[0].steps_per_launch == 1;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += align256(bytes); return at; };
  L.off_desc = take((size_t)n * sizeof(BatchDesc));
  L.off_units = take((size_t)(n + 1) * sizeof(int));
  L.off_counters = take((size_t)n * sizeof(unsigned));
  L.unit_start.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    const gbw_acoustic2d_problem& p = probs[L.order[k]];
    if (p.checkpoint_every != 0 || p.snapshot_every != 0)
      return fail(GBW_ERR_UNSUPPORTED,
                  "batched forward: checkpoints and snapshots are not supported (problem " +
                      std::to_string(L.order[k]) + ")");
    if (p.src_points > 1)
      return fail(GBW_ERR_UNSUPPORTED,
                  "batched forward: one source point per shot (problem " +
                      std::to_string(L.order[k]) + ")");
    if ((p.steps_per_launch == 1) != (L.single != 0))
      return fail(GBW_ERR_INVALID, "batched forward: steps_per_launch must agree across problems");
    Plan pl;
    int rc = make_plan(&p, 0, pl);
    if (rc) return rc;
    // default: all shots of a problem in one workgroup (the batch supplies the parallelism)
    int spb = p.shots_per_block > 0 ? p.shots_per_block : pl.S;
    spb = std::min(std::min(spb, pl.S), MAX_SPB);
    pl.spb = spb;
    pl.groups = (pl.S + spb - 1) / spb;
    if (pl.TZ2 != (probs[0].tile_rows == 24 ? 24 : 16))
      return fail(GBW_ERR_INVALID, "batched forward: tile_rows must agree across problems");
    const int tiles2 = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
    L.unit_start[k + 1] = L.unit_start[k] + tiles2 * pl.groups;
    L.plans.push_back(pl);
    L.off_prob.push_back(take(pl.total));
    L.max_nt = std::max(L.max_nt, pl.nt);
  }
  if (L.unit_start[n] <= 0) return fail(GBW_ERR_INVALID, "batch has no work");
  L.total = o;
  return 0;
}

// ---------------------------------------------------------------------------
// Batched gradient: layout and schedule (host)
// ---------------------------------------------------------------------------
struct GradLayout {
  int n = 0, T = 0, k = 0, nseg = 0, TZ2 = 16;
  std::vector<int> order;        // sorted position -> caller's problem index
  std::vector<Plan> plans;       // per sorted position
  std::vector<size_t> off_prob;  // sub-workspace of each problem
  std::vector<size_t> off_D;     // its segment bracket store (k steps), within the workspace
  std::vector<size_t> off_slots; // its checkpoint slots c0 .. nseg-1
  std::vector<int> off, c0;      // global first step, first segment
  std::vector<int> unit_start;   // n + 1 prefix sums of (two-step tiles x shot groups)
  size_t off_desc = 0, off_units = 0, total = 0;
  // problems running at global step g (sorted by first step, so a prefix)
  int active(int g) const {
    return (int)(std::upper_bound(off.begin(), off.end(), g) - off.begin());
  }
};

// default interval: about sqrt(3 T) (checkpoints 12 B/cell each, T/k of them,
// against the bracket store's 4 B/cell per step, k of them), even
int grad_auto_interval(int T) {
  int k = (int)std::lround(std::sqrt(3.0 * T) / 2.0) * 2;
  return std::max(2, std::min(k, T));
}

int make_grad_batch(const gbw_acoustic2d_problem* probs, int n, GradLayout& L) {
  if (!probs || n <= 0) return fail(GBW_ERR_INVALID, "batched gradient needs n > 0 problems");
  L = GradLayout();
  L.n = n;
  L.order.resize(n);
  for (int i = 0; i < n; ++i) L.order[i] = i;
  std::stable_sort(L.order.begin(), L.order.end(),
                   [&](int x, int y) { return probs[x].nt > probs[y].nt; });
  int max_nt = 0;
  for (int i = 0; i < n; ++i) {
    if (probs[i].nt <= 0) return fail(GBW_ERR_INVALID, "nt must be positive");
    max_nt = std::max(max_nt, probs[i].nt);
  }
  L.T = max_nt + (max_nt & 1);
  const int kreq = probs[0].checkpoint_every;
  for (int i = 1; i < n; ++i)
    if (probs[i].checkpoint_every != kreq)
      return fail(GBW_ERR_INVALID, "batched gradient: checkpoint_every must agree across problems");
  if (kreq < 0 || (kreq & 1))
    return fail(GBW_ERR_INVALID, "batched gradient: checkpoint_every must be even (0 = automatic)");
  L.k = kreq > 0 ? std::min(kreq, L.T) : grad_auto_interval(L.T);
  L.nseg = (L.T + L.k - 1) / L.k;
  L.TZ2 = probs[0].tile_rows == 24 ? 24 : 16;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += align256(bytes); return at; };
  L.off_desc = take((size_t)n * sizeof(GradDesc));
  L.off_units = take((size_t)(n + 1) * sizeof(int));
  L.unit_start.assign(n + 1, 0);
  for (int j = 0; j < n; ++j) {
    const gbw_acoustic2d_problem& p = probs[L.order[j]];
    const std::string where = "batched gradient, problem " + std::to_string(L.order[j]) + ": ";
    if (p.snapshot_every != 0 || p.checkpoint_bf16 != 0)
      return fail(GBW_ERR_UNSUPPORTED, where + "snapshots and bf16 checkpoints are not supported");
    if (p.src_points > 1)
      return fail(GBW_ERR_UNSUPPORTED, where + "one source point per shot");
    if (p.steps_per_launch == 1)
      return fail(GBW_ERR_UNSUPPORTED, where + "runs two time steps per launch (steps_per_launch 0 or 2)");
    if ((p.tile_rows == 24 ? 24 : 16) != L.TZ2)
      return fail(GBW_ERR_INVALID, where + "tile_rows must agree across problems");
    // the plan of a single gradient call with this interval (the bracket store
    // holds k steps even when the problem is shorter: it is indexed by global step)
    // (default: all shots of a problem in one workgroup, the batch supplies the parallelism)
    gbw_acoustic2d_problem q = p;
    q.checkpoint_every = L.k;
    q.nt = std::max(p.nt, L.k);
    q.shots_per_block = p.shots_per_block > 0 ? p.shots_per_block : std::max(1, p.src_n);
    Plan pl;
    int rc = make_plan(&q, 1, pl);
    if (rc) {
      g_last_error = where + g_last_error;
      return rc;
    }
    pl.nt = p.nt;
    const int tiles2 = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
    L.unit_start[j + 1] = L.unit_start[j] + tiles2 * pl.groups;
    const int off = L.T - p.nt;
    L.off.push_back(off);
    L.c0.push_back(off / L.k);
    L.off_prob.push_back(take(pl.total));
    L.off_D.push_back(pl.off_D);  // make_plan sized it for k steps (nt >= k above)
    L.off_slots.push_back(take((size_t)(L.nseg - off / L.k) * pl.slot_bytes));
    L.plans.push_back(pl);
  }
  L.total = o;
  return 0;
}

unsigned grad_prep_blocks(const GradLayout& L) {
  long long most = 0;
  for (const Plan& pl : L.plans) most = std::max(most, 3LL * pl.S * pl.d.plane);
  return (unsigned)std::max<long long>(1, std::min<long long>((most + 255) / 256, 2048));
}

template <bool RECOMP>
int grad_batch_step(const GradLayout& L, const GradDesc* descs, const int* units, int n_active,
                    int g, int j, int c, hipStream_t st) {
  const unsigned gx = (unsigned)((L.unit_start[n_active] + 7) / 8 * 8);
  LaunchTimer timer(RECOMP ? "grad_batch_step_kernel<recompute>" : "grad_batch_step_kernel<forward>", st);
  if (L.TZ2 == 24)
    grad_batch_step_kernel<24, 96, RECOMP><<<gx, FWD2_THREADS, 0, st>>>(descs, units, n_active, g, j, c, L.k);
  else
    grad_batch_step_kernel<16, 128, RECOMP><<<gx, FWD2_THREADS, 0, st>>>(descs, units, n_active, g, j, c, L.k);
  LAUNCH_CHECK();
  return 0;
}

}  // namespace

// Workspace of gbw_acoustic2d_forward_noalloc: one array per GPU, in the
// library's code object.  In host code `scratch` names a host-side stand-in,
// not device memory: its device address is looked up per device with
// hipGetSymbolAddress and only that address is passed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[GBW_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // device ordinal -> address of `scratch` there

// address of the current device's `scratch`, looked up once per device with
// that device current (each GPU holds its own copy of the array)
int scratch_address(void** ptr) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(GBW_ERR_HIP, "hipGetSymbolAddress returned no address for scratch");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}

// suppresses the timing hook's event creation for the calls in its scope
struct NoTimingScope {
  bool prev;
  NoTimingScope() : prev(t_no_timing) { t_no_timing = true; }
  ~NoTimingScope() { t_no_timing = prev; }
};

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
  const void* geo = nullptr;
  if ((rc = resolve_geometry(pl, in, workspace, st, &geo))) return rc;
  if ((rc = build_coefs(pl, prob, in, workspace, st))) return rc;
  const Dims& d = pl.d;
  const long long fstride = (long long)pl.S * d.plane;
  const long long inner = (long long)d.nz * d.nx;
  const bool energy = out->energy_p || out->energy_vx || out->energy_vz;
  float* A = at<float>(workspace, pl.off_A);
  float* B = at<float>(workspace, pl.off_B);
  float* E = at<float>(workspace, pl.off_E);
  char* ck = static_cast<char*>(out->checkpoints);
  unsigned* MX = at<unsigned>(workspace, pl.off_MX);
  const long long slot = 3 * fstride;
  if (energy) {
    if ((rc = fill(E, (long long)pl.eparts * 3 * inner, 0.f, st))) return rc;
    if (out->energy_p && (rc = fill(out->energy_p, inner, 0.f, st))) return rc;
    if (out->energy_vx && (rc = fill(out->energy_vx, inner, 0.f, st))) return rc;
    if (out->energy_vz && (rc = fill(out->energy_vz, inner, 0.f, st))) return rc;
  }
  // fp32 checkpoints are written in place of the ping-pong output (no extra
  // traffic); bf16 checkpoints are quantised from it after the step
  const bool ck32 = pl.k && !pl.bf16;
  float* cur = ck32 ? reinterpret_cast<float*>(ck) : A;
  if ((rc = fill(cur, slot, 0.f, st))) return rc;
  if (pl.k && pl.bf16 && (rc = quantize_state(pl, cur, ck, MX, st))) return rc;
  FwdArgs a = base_fwd_args(pl, prob, in, workspace, geo);
  a.rec_p = out->rec_p; a.rec_vx = out->rec_vx; a.rec_vz = out->rec_vz;
  if (!rec) a.rec_p = a.rec_vx = a.rec_vz = nullptr;
  a.energy = energy ? E : nullptr;
  const int se = prob->snapshot_every;
  int snap = 0;
  // the state after step `t` must be materialised (checkpoint, energy chunk
  // end, snapshot): a two-step launch never ends its first step there
  auto boundary = [&](int t) {
    return (pl.k && (t + 1) % pl.k == 0) || (t + 1) % pl.echunk == 0 ||
           (se > 0 && (t + 1) % se == 0);
  };
  for (int it0 = 0; it0 < pl.nt;) {
    const int steps = (pl.two && it0 + 1 < pl.nt && !boundary(it0)) ? 2 : 1;
    const int it = it0 + steps - 1;  // last step of this launch
    float* next;
    const bool ck_step = pl.k && (it + 1) % pl.k == 0 && it + 1 < pl.nt;
    if (ck_step && ck32)
      next = reinterpret_cast<float*>(ck + (size_t)((it + 1) / pl.k) * pl.slot_bytes);
    else
      next = (cur == A) ? B : A;
    set_state(a, cur, next, fstride);
    a.it = it0;
    launch_steps<false>(pl, a, steps, st);
    LAUNCH_CHECK();
    it0 += steps;
    cur = next;
    if (ck_step && pl.bf16 &&
        (rc = quantize_state(pl, cur, ck + (size_t)((it + 1) / pl.k) * pl.slot_bytes, MX, st)))
      return rc;
    const bool chunk_end = ((it + 1) % pl.echunk == 0) || it + 1 == pl.nt;
    if (energy && chunk_end) {
      energy_flush_kernel<<<(unsigned)((3 * inner + 255) / 256), 256, 0, st>>>(
          E, pl.eparts, inner, out->energy_p, out->energy_vx, out->energy_vz);
      LAUNCH_CHECK();
    }
    if (se > 0 && (((it + 1) % se == 0) || it + 1 == pl.nt)) {
      dim3 g((d.nx + 255) / 256, d.nz);
      snapshot_kernel<<<g, 256, 0, st>>>(d, cur, out->snapshots + (long long)snap * inner);
      LAUNCH_CHECK();
      ++snap;
    }
  }
  return flush_timing(st);
}

size_t gbw_acoustic2d_geometry_bytes(const gbw_acoustic2d_problem* prob) {
  Plan pl;
  if (make_plan(prob, 0, pl)) return 0;
  return pl.geo_total;
}

int gbw_acoustic2d_geometry_setup(const gbw_acoustic2d_problem* prob,
                                  const gbw_acoustic2d_inputs* in, void* geometry,
                                  size_t geometry_bytes, void* stream) {
  Plan pl;
  int rc = make_plan(prob, 0, pl);
  if (rc) return rc;
  if (!in) return fail(GBW_ERR_INVALID, "inputs is NULL");
  if (!geometry || geometry_bytes < pl.geo_total)
    return fail(GBW_ERR_WORKSPACE,
                "geometry buffer too small: need " + std::to_string(pl.geo_total) + " bytes");
  {  // a failed setup leaves no stale entry behind
    std::lock_guard<std::mutex> lk(g_geo_mu);
    g_geo.erase(geometry);
  }
  std::vector<int> rcell;
  if ((rc = setup_geometry(pl, in, geometry, static_cast<hipStream_t>(stream), rcell))) return rc;
  std::lock_guard<std::mutex> lk(g_geo_mu);
  g_geo[geometry] = geo_key(pl);
  return 0;
}

int gbw_acoustic2d_forward_noalloc(const gbw_acoustic2d_problem* prob,
                                   const gbw_acoustic2d_inputs* in,
                                   const gbw_acoustic2d_outputs* out, void* stream) {
  Plan pl;
  int rc = make_plan(prob, 0, pl);
  if (rc) return rc;
  if (pl.total > GBW_NOALLOC_SCRATCH_BYTES)
    return fail(GBW_ERR_WORKSPACE,
                "problem too large for the no-allocation forward: needs " +
                    std::to_string(pl.total) + " bytes of workspace, its scratch array holds " +
                    std::to_string((size_t)GBW_NOALLOC_SCRATCH_BYTES));
  void* ws = nullptr;
  if ((rc = scratch_address(&ws))) return rc;
  // the forward clears or writes every part of the workspace it reads (state
  // fills, receiver map, coefficients, energy partials) before reading it, so
  // what an earlier call left in the array does not matter
  NoTimingScope no_timing;
  return gbw_acoustic2d_forward(prob, in, out, ws, GBW_NOALLOC_SCRATCH_BYTES, stream);
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
  const void* geo = nullptr;
  if ((rc = resolve_geometry(pl, in, workspace, st, &geo))) return rc;
  void* geom = const_cast<void*>(geo);
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
  const char* ck = static_cast<const char*>(adj->checkpoints);
  if ((rc = fill(G1, (long long)pl.groups * d.plane, 0.f, st))) return rc;
  if (G2 && (rc = fill(G2, (long long)pl.groups * d.plane, 0.f, st))) return rc;
  if ((rc = fill(GW, (long long)pl.S * pl.K * pl.nt, 0.f, st))) return rc;
  if ((rc = fill(LA, slot, 0.f, st))) return rc;
  const bool any_g = adj->g_p || adj->g_vx || adj->g_vz;
  if (any_g && pl.R > 0) {
    adj_inject_kernel<<<(pl.S * pl.R + 255) / 256, 256, 0, st>>>(
        d, pl.S, pl.nt, pl.nt - 1, pl.R, at<int>(geom, pl.off_rcell),
        at<int>(geom, pl.off_rhead), at<int>(geom, pl.off_rnext), adj->g_p, adj->g_vx,
        adj->g_vz, LA, LA + fstride, LA + 2 * fstride);
    LAUNCH_CHECK();
  }
  FwdArgs f = base_fwd_args(pl, prob, in, workspace, geo);
  AdjArgs b{};
  b.d = d; b.S = pl.S; b.spb = pl.spb; b.tiles_x = pl.tiles_x; b.ntiles = f.ntiles;
  b.coef = f.coef; b.src_cell = f.src_cell; b.src_k = pl.K; b.nt = pl.nt; b.dt = prob->dt;
  b.rhead = f.rhead; b.rnext = f.rnext; b.rrow = f.rrow; b.R = pl.R;
  b.g_p = adj->g_p; b.g_vx = adj->g_vx; b.g_vz = adj->g_vz;
  b.gA1 = G1; b.gA2 = G2; b.gw = GW;
  float* lcur = LA;
  for (int c = pl.nseg - 1; c >= 0; --c) {
    const int t0 = c * pl.k, t1 = std::min(pl.nt, t0 + pl.k);
    // recompute the interval from its checkpoint, keeping D^n (and vx/vz brackets)
    float* cur;
    if (pl.bf16) {
      if ((rc = dequantize_state(pl, ck + (size_t)c * pl.slot_bytes, A, st))) return rc;
      cur = A;
    } else {
      cur = const_cast<float*>(reinterpret_cast<const float*>(ck + (size_t)c * pl.slot_bytes));
    }
    for (int it = t0; it < t1;) {
      const int steps = (pl.two && !pl.march && it + 1 < t1) ? 2 : 1;
      float* next = (cur == A) ? B : A;
      set_state(f, cur, next, fstride);
      f.it = it;
      f.dbuf = D + (long long)(it - t0) * fstride;
      f.gxbuf = GX ? GX + (long long)(it - t0) * fstride : nullptr;
      f.gzbuf = GZ ? GZ + (long long)(it - t0) * fstride : nullptr;
      f.bstride = fstride;
      it += steps;
      if (pl.march) {
        FwdArgs fm = f;
        fm.ntiles = pl.tiles_x * pl.mseg * pl.S;
        fm.seg_h = pl.mseg_h;
        LaunchTimer timer("recomp_march_kernel", st);
        recomp_march_kernel<<<(unsigned)((fm.ntiles + 7) / 8 * 8), NTHREADS, 0, st>>>(fm);
      } else {
        launch_steps<true>(pl, f, steps, st);
      }
      LAUNCH_CHECK();
      cur = next;
    }
    for (int it = t1 - 1; it >= t0;) {
      const int steps = (pl.two && it - 1 >= t0) ? 2 : 1;
      float* lnext = (lcur == LA) ? LB : LA;
      b.LP0 = lcur; b.LX0 = lcur + fstride; b.LZ0 = lcur + 2 * fstride;
      b.LP1 = lnext; b.LX1 = lnext + fstride; b.LZ1 = lnext + 2 * fstride;
      b.it = it;
      b.dbuf = D + (long long)(it - t0) * fstride;
      b.gxbuf = GX ? GX + (long long)(it - t0) * fstride : nullptr;
      b.gzbuf = GZ ? GZ + (long long)(it - t0) * fstride : nullptr;
      if (steps == 2) {
        b.dbuf1 = b.dbuf - fstride;
        b.gxbuf1 = GX ? b.gxbuf - fstride : nullptr;
        b.gzbuf1 = GZ ? b.gzbuf - fstride : nullptr;
        launch_adj2(pl, b, st);
      } else if (pl.TZ == 32) {
        launch_adj<32>(pl, b, st);
      } else {
        launch_adj<16>(pl, b, st);
      }
      LAUNCH_CHECK();
      lcur = lnext;
      it -= steps;
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
  if (adj->grad_src_v && pl.K > 1 && prob->wavelet_shared == 0) {
    const int n = pl.S * pl.nt;
    wavelet_grad_points_kernel<<<(n + 255) / 256, 256, 0, st>>>(GW, pl.S, pl.K, pl.nt,
                                                                adj->grad_src_v);
    LAUNCH_CHECK();
  } else if (adj->grad_src_v) {
    // shared: summed over shots; per shot or per point: the samples as they are
    const int shared = prob->wavelet_shared == 1;
    const int n = shared ? pl.nt : pl.S * pl.K * pl.nt;
    wavelet_grad_kernel<<<(n + 255) / 256, 256, 0, st>>>(GW, pl.S * pl.K, pl.nt, shared,
                                                         adj->grad_src_v);
    LAUNCH_CHECK();
  }
  return flush_timing(st);
}

size_t gbw_acoustic2d_batch_workspace_bytes(const gbw_acoustic2d_problem* probs, int n) {
  BatchLayout L;
  if (make_batch(probs, n, L)) return 0;
  return L.total;
}

int gbw_acoustic2d_batch_setup(const gbw_acoustic2d_problem* probs,
                               const gbw_acoustic2d_inputs* ins,
                               const gbw_acoustic2d_outputs* outs, int n, void* workspace,
                               size_t workspace_bytes, void* stream) {
  BatchLayout L;
  int rc = make_batch(probs, n, L);
  if (rc) return rc;
  if (!ins || !outs) return fail(GBW_ERR_INVALID, "inputs / outputs are NULL");
  if (!workspace || workspace_bytes < L.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<BatchDesc> descs(n);
  for (int k = 0; k < n; ++k) {
    const int o = L.order[k];
    const Plan& pl = L.plans[k];
    void* wsk = static_cast<char*>(workspace) + L.off_prob[k];
    const gbw_acoustic2d_outputs& out = outs[o];
    if (out.energy_p || out.energy_vx || out.energy_vz || out.snapshots || out.checkpoints)
      return fail(GBW_ERR_UNSUPPORTED, "batched forward returns seismograms only (problem " +
                                           std::to_string(o) + ")");
    const bool rec = out.rec_p || out.rec_vx || out.rec_vz;
    if (rec && !(out.rec_p && out.rec_vx && out.rec_vz))
      return fail(GBW_ERR_INVALID, "rec_p, rec_vx and rec_vz must be given together");
    std::vector<int> rcell;
    void* geo = static_cast<char*>(wsk) + pl.off_geo;
    if ((rc = setup_geometry(pl, &ins[o], geo, st, rcell))) {
      g_last_error = "problem " + std::to_string(o) + ": " + g_last_error;
      return rc;
    }
    if ((rc = build_coefs(pl, &probs[o], &ins[o], wsk, st))) return rc;
    BatchDesc& b = descs[k];
    b = BatchDesc{};
    b.a = base_fwd_args(pl, &probs[o], &ins[o], wsk, geo);
    b.a.spb = pl.spb;
    b.a.tiles_x = pl.tiles_x2;
    b.a.ntiles = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
    b.a.rec_p = rec ? out.rec_p : nullptr;
    b.a.rec_vx = rec ? out.rec_vx : nullptr;
    b.a.rec_vz = rec ? out.rec_vz : nullptr;
    b.a.energy = nullptr;
    b.A = at<float>(wsk, pl.off_A);
    b.B = at<float>(wsk, pl.off_B);
    b.fstride = (long long)pl.S * pl.d.plane;
    b.groups = pl.groups;
  }
  HIP_TRY(hipMemcpyAsync(at<BatchDesc>(workspace, L.off_desc), descs.data(),
                         descs.size() * sizeof(BatchDesc), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at<int>(workspace, L.off_units), L.unit_start.data(),
                         L.unit_start.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));  // the host vectors go out of scope
  return 0;
}

int gbw_acoustic2d_batch_forward(const gbw_acoustic2d_problem* probs, int n, void* workspace,
                                 size_t workspace_bytes, unsigned long long* done_ticks,
                                 void* stream) {
  BatchLayout L;
  int rc = make_batch(probs, n, L);
  if (rc) return rc;
  if (!workspace || workspace_bytes < L.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const BatchDesc* descs = at<BatchDesc>(workspace, L.off_desc);
  const int* units = at<int>(workspace, L.off_units);
  unsigned* counters = at<unsigned>(workspace, L.off_counters);
  if (done_ticks) {
    batch_start_kernel<<<1, 64, 0, st>>>(done_ticks);
    LAUNCH_CHECK();
    fill_int_kernel<<<1, 256, 0, st>>>(reinterpret_cast<int*>(counters), n, 0);
    LAUNCH_CHECK();
  }
  batch_zero_kernel<<<dim3(256, n), 256, 0, st>>>(descs);
  LAUNCH_CHECK();
  const int step = L.single ? 1 : 2;
  int n_active = n, parity = 0;
  for (int it = 0; it < L.max_nt; it += step) {
    while (n_active > 0 && L.plans[n_active - 1].nt <= it) --n_active;
    const int nunits = L.unit_start[n_active];
    LaunchTimer timer("fwd2_batch_kernel", st);
    const unsigned gx = (unsigned)((nunits + 7) / 8 * 8);
    if (L.plans[0].TZ2 == 24)
      fwd2_batch_kernel<24, 96><<<gx, FWD2_THREADS, 0, st>>>(descs, units, n_active, it, parity,
                                                            L.single, counters, done_ticks);
    else
      fwd2_batch_kernel<16, 128><<<gx, FWD2_THREADS, 0, st>>>(descs, units, n_active, it, parity,
                                                             L.single, counters, done_ticks);
    LAUNCH_CHECK();
    parity ^= 1;
  }
  return flush_timing(st);
}

size_t gbw_acoustic2d_batch_gradient_workspace_bytes(const gbw_acoustic2d_problem* probs, int n) {
  GradLayout L;
  if (make_grad_batch(probs, n, L)) return 0;
  return L.total;
}

int gbw_acoustic2d_batch_gradient_setup(const gbw_acoustic2d_problem* probs,
                                        const gbw_acoustic2d_inputs* ins,
                                        const gbw_acoustic2d_outputs* outs,
                                        const gbw_acoustic2d_adjoint* adjs, int n,
                                        void* workspace, size_t workspace_bytes, void* stream) {
  GradLayout L;
  int rc = make_grad_batch(probs, n, L);
  if (rc) return rc;
  if (!ins || !outs || !adjs) return fail(GBW_ERR_INVALID, "inputs / outputs / adjoints are NULL");
  if (!workspace || workspace_bytes < L.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<GradDesc> descs(n);
  for (int j = 0; j < n; ++j) {
    const int o = L.order[j];
    const Plan& pl = L.plans[j];
    void* wsk = static_cast<char*>(workspace) + L.off_prob[j];
    const gbw_acoustic2d_outputs& out = outs[o];
    const gbw_acoustic2d_adjoint& adj = adjs[o];
    const std::string where = "problem " + std::to_string(o) + ": ";
    if (out.energy_p || out.energy_vx || out.energy_vz || out.snapshots)
      return fail(GBW_ERR_UNSUPPORTED, "batched gradient returns seismograms only (" + where + ")");
    const bool rec = out.rec_p || out.rec_vx || out.rec_vz;
    if (rec && !(out.rec_p && out.rec_vx && out.rec_vz))
      return fail(GBW_ERR_INVALID, where + "rec_p, rec_vx and rec_vz must be given together");
    if (adj.grad_rho)
      return fail(GBW_ERR_UNSUPPORTED, where + "the batched gradient computes vp and src_v gradients");
    std::vector<int> rcell;
    void* geo = static_cast<char*>(wsk) + pl.off_geo;
    if ((rc = setup_geometry(pl, &ins[o], geo, st, rcell))) {
      g_last_error = where + g_last_error;
      return rc;
    }
    if ((rc = build_coefs(pl, &probs[o], &ins[o], wsk, st))) return rc;
    GradDesc& g = descs[j];
    g = GradDesc{};
    g.f = base_fwd_args(pl, &probs[o], &ins[o], wsk, geo);
    g.f.spb = pl.spb;
    g.f.tiles_x = pl.tiles_x2;
    g.f.ntiles = pl.tiles_x2 * ((pl.d.NZ + pl.TZ2 - 1) / pl.TZ2);
    g.f.rec_p = rec ? out.rec_p : nullptr;
    g.f.rec_vx = rec ? out.rec_vx : nullptr;
    g.f.rec_vz = rec ? out.rec_vz : nullptr;
    g.f.energy = nullptr;
    AdjArgs& b = g.b;
    b.d = pl.d; b.S = pl.S; b.spb = pl.spb; b.tiles_x = g.f.tiles_x; b.ntiles = g.f.ntiles;
    b.coef = g.f.coef; b.src_cell = g.f.src_cell; b.nt = pl.nt; b.dt = probs[o].dt;
    b.rhead = g.f.rhead; b.rnext = g.f.rnext; b.rrow = g.f.rrow; b.R = pl.R;
    b.g_p = adj.g_p; b.g_vx = adj.g_vx; b.g_vz = adj.g_vz;
    g.A = at<float>(wsk, pl.off_A);
    g.B = at<float>(wsk, pl.off_B);
    g.LA = at<float>(wsk, pl.off_LA);
    g.LB = at<float>(wsk, pl.off_LB);
    g.D = at<float>(wsk, L.off_D[j]);
    g.slots = static_cast<char*>(workspace) + L.off_slots[j];
    g.G1 = at<float>(wsk, pl.off_G1);
    g.GW = at<float>(wsk, pl.off_GW);
    b.gA1 = g.G1; b.gA2 = nullptr; b.gw = g.GW;
    g.rcell = at<int>(geo, pl.off_rcell);
    g.vp = ins[o].vp; g.rho = ins[o].rho;
    g.grad_vp = adj.grad_vp; g.grad_w = adj.grad_src_v;
    g.dz = probs[o].dz;
    g.shared = probs[o].wavelet_shared == 1;
    g.fstride = (long long)pl.S * pl.d.plane;
    g.slot_bytes = (long long)pl.slot_bytes;
    g.off = L.off[j];
    g.c0 = L.c0[j];
    g.groups = pl.groups;
  }
  HIP_TRY(hipMemcpyAsync(at<GradDesc>(workspace, L.off_desc), descs.data(),
                         descs.size() * sizeof(GradDesc), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at<int>(workspace, L.off_units), L.unit_start.data(),
                         L.unit_start.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));  // the host vectors go out of scope
  return 0;
}

int gbw_acoustic2d_batch_gradient_forward(const gbw_acoustic2d_problem* probs, int n,
                                          void* workspace, size_t workspace_bytes,
                                          unsigned long long* done_ticks, void* stream) {
  GradLayout L;
  int rc = make_grad_batch(probs, n, L);
  if (rc) return rc;
  if (!workspace || workspace_bytes < L.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const GradDesc* descs = at<GradDesc>(workspace, L.off_desc);
  const int* units = at<int>(workspace, L.off_units);
  if (done_ticks) {
    stamp_kernel<<<1, 64, 0, st>>>(done_ticks);
    LAUNCH_CHECK();
  }
  const unsigned pb = grad_prep_blocks(L);
  grad_batch_prep_kernel<<<dim3(pb, n), 256, 0, st>>>(descs, 0, 0, 0);
  LAUNCH_CHECK();
  for (int g = 0, j = 0; g < L.T; g += 2, ++j) {
    const int na = L.active(g + 1);
    if ((rc = grad_batch_step<false>(L, descs, units, na, g, j, 0, st))) return rc;
    if ((g + 2) % L.k == 0 && g + 2 < L.T) {  // checkpoint: launch j wrote A (j odd) or B
      grad_batch_prep_kernel<<<dim3(pb, na), 256, 0, st>>>(descs, 3, (g + 2) / L.k, (j & 1) ? 0 : 1);
      LAUNCH_CHECK();
    }
  }
  return flush_timing(st);
}

int gbw_acoustic2d_batch_gradient_backward(const gbw_acoustic2d_problem* probs, int n,
                                           void* workspace, size_t workspace_bytes,
                                           unsigned long long* done_ticks, void* stream) {
  GradLayout L;
  int rc = make_grad_batch(probs, n, L);
  if (rc) return rc;
  if (!workspace || workspace_bytes < L.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(L.total) + " bytes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const GradDesc* descs = at<GradDesc>(workspace, L.off_desc);
  const int* units = at<int>(workspace, L.off_units);
  const unsigned pb = grad_prep_blocks(L);
  grad_batch_prep_kernel<<<dim3(pb, n), 256, 0, st>>>(descs, 2, 0, 0);
  LAUNCH_CHECK();
  int most_sr = 1;
  for (const Plan& pl : L.plans) most_sr = std::max(most_sr, pl.S * pl.R);
  grad_batch_inject_kernel<<<dim3((most_sr + 255) / 256, n), 256, 0, st>>>(descs);
  LAUNCH_CHECK();
  int parity = 0;  // the adjoint state at the last step is in LA
  for (int c = L.nseg - 1; c >= 0; --c) {
    const int lo = c * L.k, hi = std::min(L.T, lo + L.k);
    const int n_seg = L.active(hi - 1);
    grad_batch_prep_kernel<<<dim3(pb, n_seg), 256, 0, st>>>(descs, 1, c, 0);
    LAUNCH_CHECK();
    for (int g = lo, j = 0; g < hi; g += 2, ++j)
      if ((rc = grad_batch_step<true>(L, descs, units, L.active(g + 1), g, j, c, st))) return rc;
    for (int g = hi - 1; g > lo; g -= 2) {
      const int na = L.active(g);
      {
        const unsigned gx = (unsigned)((L.unit_start[na] + 7) / 8 * 8);
        AdjArgs ba{};
        ba.batch_descs = descs; ba.batch_units = units; ba.batch_n = na; ba.batch_g = g;
        ba.batch_parity = parity; ba.batch_c = c; ba.batch_k = L.k;
        LaunchTimer timer("adj2_step_kernel<batch>", st);
        if (L.TZ2 == 24)
          adj2_step_kernel<24, 96, FWD2_THREADS, false, true><<<gx, FWD2_THREADS, 0, st>>>(ba);
        else
          adj2_step_kernel<16, 128, FWD2_THREADS, false, true><<<gx, FWD2_THREADS, 0, st>>>(ba);
        LAUNCH_CHECK();
      }
      parity ^= 1;
      // problems whose first step was g or g-1 are done: chain rule, wavelet gradient, stamp
      for (int j = L.active(g - 2); j < na; ++j) {
        const Plan& pl = L.plans[j];
        grad_batch_fold_kernel<<<dim3((pl.d.nx + 127) / 128, pl.d.nz), 128, 0, st>>>(descs, j);
        LAUNCH_CHECK();
        const int nw = pl.S * pl.nt;
        grad_batch_wgrad_kernel<<<(nw + 255) / 256, 256, 0, st>>>(descs, j);
        LAUNCH_CHECK();
        if (done_ticks) {
          stamp_kernel<<<1, 64, 0, st>>>(done_ticks + 1 + j);
          LAUNCH_CHECK();
        }
      }
    }
  }
  return flush_timing(st);
}

int gbw_device_wall_clock_khz(void) {
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0;
  return khz;
}

void gbw_set_kernel_timing(int on) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing = on != 0;
  g_times.clear();
}

size_t gbw_kernel_times(char* buf, size_t len) {
  std::string out;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    for (const auto& kv : g_times)
      out += kv.first + " " + std::to_string(kv.second.first) + " " +
             std::to_string(kv.second.second) + "\n";
  }
  if (buf && len > 0) {
    const size_t n = std::min(len - 1, out.size());
    std::memcpy(buf, out.data(), n);
    buf[n] = '\0';
  }
  return out.size() + 1;
}

}  // extern "C"
