// gbw_acoustic3d.hip -- MI355X (gfx950) kernels of the 3D acoustic propagator.
//
// Replaces the time stepping of GeoBrain's acoustic_forward_kernel_3d
// (geobrain/physics/wave/kernels/acoustic_3d.py:20-476): the 4th-order
// staggered-grid pressure / velocity scheme with split PML damping, point
// sources, receiver sampling and energy maps, forward only.
//
// Layout in HBM: each field (p, vx, vy, vz) is [shot][NZ][NY][PX] float32 with
// PX = NX rounded up to a multiple of 64, so every (z, y) row starts on a
// 256-byte boundary; the padding columns, vx's column NX-1, vy's row NY-1 and
// vz's plane NZ-1 stay zero.  The fields are updated in place: the p-update of
// a cell reads its own p and neighbouring velocities only, the v-update its
// own velocities and neighbouring p only.
//
// Per time step, three launches:
//   p3_step_kernel   p on its region, then the point source (acoustic_3d.py:124-169)
//   v3_step_kernel   the mirror p[f-1] = -p[f+1], vx / vy / vz on their regions,
//                    vz[f-1] = vz[f], and the energy sums of the step (:179-262)
//   record3_kernel   p, vx, vy, vz at the receivers (:244-248)
// A thread owns four consecutive cells of a row (one float4 group) and steps
// them through the shots of its shot group, with the coefficients in
// registers.  Arithmetic rounds once per operation in the reference's order
// (-ffp-contract=off), so the seismograms match the reference's FP32 PyTorch
// path bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "gbw_acoustic3d.h"
#include "gbw_common.h"

using namespace gbw;

namespace {

constexpr float k3C1 = 9.0f / 8.0f;
constexpr float k3C2 = (float)(-1.0 / 24.0);
constexpr int T3 = 256;  // threads per workgroup

struct Dims3 {
  int nz, ny, nx, nabc, NZ, NY, NX, PX, GX, f, fs;
  long long plane;  // NZ * NY * PX
  long long groups;  // NZ * NY * GX float4 groups per plane
};

__device__ __forceinline__ float g3(const float4& v, int e) {
  return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}
__device__ __forceinline__ void s3(float4& v, int e, float s) {
  if (e == 0) v.x = s; else if (e == 1) v.y = s; else if (e == 2) v.z = s; else v.w = s;
}
// c1 * (a - b) + c2 * (c - d), one rounding per operation
__device__ __forceinline__ float d4(float a, float b, float c, float d) {
  const float t1 = k3C1 * (a - b);
  const float t2 = k3C2 * (c - d);
  return t1 + t2;
}

// workgroup index -> block of 256 groups: consecutive workgroups go to
// different XCDs (round-robin over 8), so give each XCD a contiguous range
__device__ __forceinline__ long long xcd_block3() {
  const long long b = blockIdx.x;
  return (b & 7) * (long long)(gridDim.x >> 3) + (b >> 3);
}

// coefficients (acoustic_3d.py:401-431), pad_model_3d as a clamp map
__global__ void coef3_kernel(Dims3 d, const float* __restrict__ vp, const float* __restrict__ rho,
                             const float* __restrict__ damp, float dt, float dz,
                             float* __restrict__ coef) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.plane) return;
  const int x = (int)(q % d.PX);
  const long long zy = q / d.PX;
  const int y = (int)(zy % d.NY), z = (int)(zy / d.NY);
  float omk1 = 1.f, a1 = 0.f, omk2 = 1.f, omk3 = 1.f, omk4 = 1.f, a2 = 0.f;
  if (x < d.NX) {
    const int a = min(max(z - d.nabc, 0), d.nz - 1);
    const int b = min(max(y - d.nabc, 0), d.ny - 1);
    const int e = min(max(x - d.nabc, 0), d.nx - 1);
    const long long m = ((long long)a * d.ny + b) * d.nx + e;
    const float c = vp[m], den = rho[m];
    float t = den * c;
    t = t * c;
    t = t * dt;
    a1 = t / dz;
    const long long dq = ((long long)z * d.NY + y) * d.NX + x;  // damp is (NZ, NY, NX)
    const float k1 = damp[dq] * dt;
    omk1 = 1.f - k1;
    const float dd = den * dz;
    const float r = 1.f / dd;
    a2 = r * dt;
    float k2 = 0.f, k3 = 0.f, k4 = 0.f;
    if (x >= 1 && x <= d.NX - 3) {
      float s = damp[dq] + damp[dq + 1];
      s = 0.5f * s;
      k2 = s * dt;
    }
    if (y >= 1 && y <= d.NY - 3) {
      float s = damp[dq] + damp[dq + d.NX];
      s = 0.5f * s;
      k3 = s * dt;
    }
    if (z >= d.f && z <= d.NZ - 3) {
      float s = damp[dq] + damp[dq + (long long)d.NY * d.NX];
      s = 0.5f * s;
      k4 = s * dt;
    }
    omk2 = 1.f - k2;
    omk3 = 1.f - k3;
    omk4 = 1.f - k4;
  }
  coef[q] = omk1;
  coef[d.plane + q] = a1;
  coef[2 * d.plane + q] = omk2;
  coef[3 * d.plane + q] = omk3;
  coef[4 * d.plane + q] = omk4;
  coef[5 * d.plane + q] = a2;
}

struct Step3Args {
  Dims3 d;
  int S, spb, it, nt;
  float dt;
  float* state;            // [field][shot][plane]
  long long fstride;       // S * plane
  const float* coef;
  const int* src_cell;     // [S] padded cell index in the plane
  const float* wav; long long wav_stride;
  float* energy;           // [group][4][nz][ny][nx] or null
};

// p on z in [f+1, NZ-3], y in [2, NY-3], x in [2, NX-3]; then the source
__global__ __launch_bounds__(T3) void p3_step_kernel(Step3Args a) {
  const Dims3 d = a.d;
  const long long gq = xcd_block3() * T3 + threadIdx.x;  // float4 group of the plane
  if (gq >= d.groups) return;
  const int g = (int)(gq % d.GX);
  const long long zy = gq / d.GX;
  const int y = (int)(zy % d.NY), z = (int)(zy / d.NY);
  const int s_begin = blockIdx.y * a.spb, s_end = min(a.S, s_begin + a.spb);
  const long long row = (long long)z * d.NY + y;
  const long long q4 = row * d.GX + g;  // float4 index
  unsigned mask = 0;
  if (z >= d.f + 1 && z <= d.NZ - 3 && y >= 2 && y <= d.NY - 3)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 4 * g + e;
      if (x >= 2 && x <= d.NX - 3) mask |= 1u << e;
    }
  const float4* __restrict__ C = reinterpret_cast<const float4*>(a.coef);
  const long long p4 = d.plane >> 2;
  const float4 omk1 = C[q4], a1 = C[p4 + q4];
  const long long sy4 = d.GX, sz4 = (long long)d.NY * d.GX;
  const long long gl = row * d.GX + max(g - 1, 0), gr = row * d.GX + min(g + 1, d.GX - 1);
  for (int s = s_begin; s < s_end; ++s) {
    float4* P = reinterpret_cast<float4*>(a.state + (long long)s * d.plane);
    const float4* VX = reinterpret_cast<const float4*>(a.state + a.fstride + (long long)s * d.plane);
    const float4* VY = reinterpret_cast<const float4*>(a.state + 2 * a.fstride + (long long)s * d.plane);
    const float4* VZ = reinterpret_cast<const float4*>(a.state + 3 * a.fstride + (long long)s * d.plane);
    const int sc = a.src_cell[s];
    const bool src_here = (long long)(sc >> 2) == q4;
    if (!mask && !src_here) continue;
    float4 p = P[q4];
    if (mask) {
      const float4 xa = VX[gl], xb = VX[q4], xc = VX[gr];
      const float vxs[12] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w};
      const float4 ym2 = VY[q4 - 2 * sy4], ym1 = VY[q4 - sy4], y00 = VY[q4], yp1 = VY[q4 + sy4];
      const float4 zm2 = VZ[q4 - 2 * sz4], zm1 = VZ[q4 - sz4], z00 = VZ[q4], zp1 = VZ[q4 + sz4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((mask >> e) & 1u)) continue;
        const float dvx = d4(vxs[4 + e], vxs[3 + e], vxs[5 + e], vxs[2 + e]);
        const float dvy = d4(g3(y00, e), g3(ym1, e), g3(yp1, e), g3(ym2, e));
        const float dvz = d4(g3(z00, e), g3(zm1, e), g3(zp1, e), g3(zm2, e));
        float sum = dvx + dvy;
        sum = sum + dvz;
        const float u = g3(omk1, e) * g3(p, e);
        s3(p, e, u - g3(a1, e) * sum);
      }
    }
    if (src_here) {
      const int e = sc & 3;
      s3(p, e, g3(p, e) + a.dt * a.wav[(long long)s * a.wav_stride + a.it]);
    }
    P[q4] = p;
  }
}

// the p mirror, vx / vy / vz on their regions, the vz mirror, energy sums
__global__ __launch_bounds__(T3) void v3_step_kernel(Step3Args a) {
  const Dims3 d = a.d;
  const long long gq = xcd_block3() * T3 + threadIdx.x;
  if (gq >= d.groups) return;
  const int g = (int)(gq % d.GX);
  const long long zy = gq / d.GX;
  const int y = (int)(zy % d.NY), z = (int)(zy / d.NY);
  const int s_begin = blockIdx.y * a.spb, s_end = min(a.S, s_begin + a.spb);
  const long long row = (long long)z * d.NY + y;
  const long long q4 = row * d.GX + g;
  const bool zx = z >= d.f && z <= d.NZ - 2, zz = z >= d.f && z <= d.NZ - 3;
  unsigned mx = 0, my = 0, mz = 0, me = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int x = 4 * g + e;
    if (zx && y >= 1 && y <= d.NY - 2 && x >= 1 && x <= d.NX - 3) mx |= 1u << e;
    if (zx && y >= 1 && y <= d.NY - 3 && x >= 1 && x <= d.NX - 2) my |= 1u << e;
    if (zz && y >= 1 && y <= d.NY - 2 && x >= 1 && x <= d.NX - 2) mz |= 1u << e;
    if (z >= d.nabc && z < d.nabc + d.nz && y >= d.nabc && y < d.nabc + d.ny && x >= d.nabc &&
        x < d.nabc + d.nx)
      me |= 1u << e;
  }
  const bool energy = a.energy && me;
  const bool mirror_p = d.fs && z == d.f - 1;  // p[f-1] = -p[f+1]
  const bool vz_up = d.fs && z == d.f;         // vz[f-1] = vz[f]
  if (!(mx | my | mz) && !energy && !mirror_p) return;
  const float4* __restrict__ C = reinterpret_cast<const float4*>(a.coef);
  const long long p4 = d.plane >> 2;
  const float4 omk2 = C[2 * p4 + q4], omk3 = C[3 * p4 + q4], omk4 = C[4 * p4 + q4];
  const float4 a2 = C[5 * p4 + q4];
  const long long sy4 = d.GX, sz4 = (long long)d.NY * d.GX;
  const long long gl = row * d.GX + max(g - 1, 0), gr = row * d.GX + min(g + 1, d.GX - 1);
  float4 ep = make_float4(0.f, 0.f, 0.f, 0.f), ex = ep, ey = ep, ez = ep;
  auto add_sq = [](float4& acc, const float4& v) {
    acc = make_float4(acc.x + v.x * v.x, acc.y + v.y * v.y, acc.z + v.z * v.z, acc.w + v.w * v.w);
  };
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    float4* P = reinterpret_cast<float4*>(a.state + so);
    float4* VX = reinterpret_cast<float4*>(a.state + a.fstride + so);
    float4* VY = reinterpret_cast<float4*>(a.state + 2 * a.fstride + so);
    float4* VZ = reinterpret_cast<float4*>(a.state + 3 * a.fstride + so);
    if (mirror_p) {  // no velocity of row f-1 is updated; nothing else reads this row here
      const float4 v = P[q4 + 2 * sz4];
      P[q4] = make_float4(-v.x, -v.y, -v.z, -v.w);
      continue;
    }
    const float4 pb = P[q4];
    float4 vx = VX[q4], vy = VY[q4], vz = VZ[q4];
    if (mx) {
      const float4 pa = P[gl], pc = P[gr];
      const float ps[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w, pc.x, pc.y, pc.z, pc.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((mx >> e) & 1u)) continue;
        const float u = g3(omk2, e) * g3(vx, e);
        s3(vx, e, u - g3(a2, e) * d4(ps[5 + e], ps[4 + e], ps[6 + e], ps[3 + e]));
      }
      VX[q4] = vx;
    }
    if (my) {
      const float4 pu = P[q4 - sy4], pd1 = P[q4 + sy4], pd2 = P[q4 + 2 * sy4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((my >> e) & 1u)) continue;
        const float u = g3(omk3, e) * g3(vy, e);
        s3(vy, e, u - g3(a2, e) * d4(g3(pd1, e), g3(pb, e), g3(pd2, e), g3(pu, e)));
      }
      VY[q4] = vy;
    }
    if (mz) {
      // at z = f the plane above is the mirrored row f-1: read it as -p[f+1]
      const float4 pd1 = P[q4 + sz4], pd2 = P[q4 + 2 * sz4];
      float4 pu;
      if (vz_up) {
        pu = make_float4(-pd1.x, -pd1.y, -pd1.z, -pd1.w);
      } else {
        pu = P[q4 - sz4];
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((mz >> e) & 1u)) continue;
        const float u = g3(omk4, e) * g3(vz, e);
        s3(vz, e, u - g3(a2, e) * d4(g3(pd1, e), g3(pb, e), g3(pd2, e), g3(pu, e)));
      }
      VZ[q4] = vz;
      if (vz_up) VZ[q4 - sz4] = vz;
    } else if (vz_up) {
      VZ[q4 - sz4] = vz;  // cells of row f outside the vz region: copy as is
    }
    if (energy) {
      add_sq(ep, pb);
      add_sq(ex, vx);
      add_sq(ey, vy);
      add_sq(ez, vz);
    }
  }
  if (energy) {
    const long long inner = (long long)d.nz * d.ny * d.nx;
    float* eg = a.energy + (long long)blockIdx.y * 4 * inner;
    const long long ob = ((long long)(z - d.nabc) * d.ny + (y - d.nabc)) * d.nx;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (!((me >> e) & 1u)) continue;
      const long long o = ob + (4 * g + e - d.nabc);
      eg[o] = eg[o] + g3(ep, e);
      eg[inner + o] = eg[inner + o] + g3(ex, e);
      eg[2 * inner + o] = eg[2 * inner + o] + g3(ey, e);
      eg[3 * inner + o] = eg[3 * inner + o] + g3(ez, e);
    }
  }
}

// p, vx, vy, vz at the receivers (S x R threads)
__global__ void record3_kernel(Dims3 d, int S, int R, int nt, int it, const float* state,
                               long long fstride, const int* rcell, float* rp, float* rvx,
                               float* rvy, float* rvz) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)S * R) return;
  const int s = (int)(i / R), r = (int)(i % R);
  const long long q = (long long)s * d.plane + rcell[r];
  const long long o = ((long long)s * nt + it) * R + r;
  rp[o] = state[q];
  rvx[o] = state[fstride + q];
  rvy[o] = state[2 * fstride + q];
  rvz[o] = state[3 * fstride + q];
}

// chunk end: totals += sum over shot groups of the partial sums; partials -> 0
__global__ void energy3_flush_kernel(float* part, int groups, long long inner, float* e0,
                                     float* e1, float* e2, float* e3) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * inner) return;
  const int f = (int)(i / inner);
  const long long q = i % inner;
  float acc = part[i];
  part[i] = 0.f;
  for (int g = 1; g < groups; ++g) {
    acc = acc + part[(long long)g * 4 * inner + i];
    part[(long long)g * 4 * inner + i] = 0.f;
  }
  float* out = f == 0 ? e0 : (f == 1 ? e1 : (f == 2 ? e2 : e3));
  if (out) out[q] = out[q] + acc;
}

__global__ void fill3_kernel(float* __restrict__ p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Plan3 {
  Dims3 d;
  int S, R, nt, spb, groups, echunk;
  size_t off_coef, off_state, off_E, off_src, off_rcell, total;
};

int make_plan3(const gbw_acoustic3d_problem* p, Plan3& pl) {
  if (!p) return fail(GBW_ERR_INVALID, "problem is NULL");
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
    return fail(GBW_ERR_INVALID, "Grid dimensions must be positive");
  if (p->nabc < 1) return fail(GBW_ERR_INVALID, "nabc must be >= 1");
  if (p->nt <= 0) return fail(GBW_ERR_INVALID, "nt must be positive");
  if (p->src_n <= 0) return fail(GBW_ERR_INVALID, "src_n must be positive");
  if (p->rcv_n < 0) return fail(GBW_ERR_INVALID, "rcv_n must be >= 0");
  if (!(p->dt > 0.f) || !(p->dz > 0.f)) return fail(GBW_ERR_INVALID, "dt and dz must be positive");
  Dims3& d = pl.d;
  d.nz = p->nz; d.ny = p->ny; d.nx = p->nx; d.nabc = p->nabc;
  d.NZ = p->nz + 2 * p->nabc; d.NY = p->ny + 2 * p->nabc; d.NX = p->nx + 2 * p->nabc;
  d.PX = (d.NX + 63) / 64 * 64;
  d.GX = d.PX / 4;
  d.fs = p->free_surface ? 1 : 0;
  d.f = d.fs ? p->nabc : 1;
  d.plane = (long long)d.NZ * d.NY * d.PX;
  d.groups = d.plane / 4;
  if (d.NZ < 6 || d.NY < 6 || d.NX < 6) return fail(GBW_ERR_INVALID, "padded grid must be at least 6 per axis");
  if (d.plane >= (1LL << 31)) return fail(GBW_ERR_INVALID, "padded volume exceeds 2^31 cells");
  pl.S = p->src_n; pl.R = p->rcv_n; pl.nt = p->nt;
  int spb = p->shots_per_block > 0 ? std::min(p->shots_per_block, pl.S) : pl.S;
  pl.spb = spb;
  pl.groups = (pl.S + spb - 1) / spb;
  if (pl.groups > 65535) return fail(GBW_ERR_INVALID, "more than 65535 shot groups: raise shots_per_block");
  pl.echunk = p->energy_chunk > 0 ? std::min(p->energy_chunk, p->nt) : p->nt;
  const size_t inner = (size_t)d.nz * d.ny * d.nx;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += al256(bytes); return at; };
  pl.off_coef = take(6 * (size_t)d.plane * sizeof(float));
  pl.off_state = take(4 * (size_t)pl.S * d.plane * sizeof(float));
  pl.off_E = take((size_t)pl.groups * 4 * inner * sizeof(float));
  pl.off_src = take((size_t)pl.S * sizeof(int));
  pl.off_rcell = take((size_t)std::max(1, pl.R) * sizeof(int));
  pl.total = o;
  return 0;
}

template <typename T>
T* at3(void* ws, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(ws) + off); }

int fill3(float* p, long long n, float v, hipStream_t st) {
  if (n <= 0) return 0;
  fill3_kernel<<<(unsigned)std::min<long long>((n + 255) / 256, 16384), 256, 0, st>>>(p, n, v);
  LAUNCH_CHECK();
  return 0;
}

}  // namespace

extern "C" {

size_t gbw_acoustic3d_workspace_bytes(const gbw_acoustic3d_problem* prob) {
  Plan3 pl;
  if (make_plan3(prob, pl)) return 0;
  return pl.total;
}

int gbw_acoustic3d_forward(const gbw_acoustic3d_problem* prob, const gbw_acoustic3d_inputs* in,
                           const gbw_acoustic3d_outputs* out, void* workspace,
                           size_t workspace_bytes, void* stream) {
  Plan3 pl;
  int rc = make_plan3(prob, pl);
  if (rc) return rc;
  if (!in || !out) return fail(GBW_ERR_INVALID, "inputs / outputs are NULL");
  if (!workspace || workspace_bytes < pl.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(pl.total) + " bytes");
  if (!in->vp || !in->rho || !in->damp || !in->src_v)
    return fail(GBW_ERR_INVALID, "vp, rho, damp and src_v must be device pointers");
  if (!in->src_x || !in->src_y || !in->src_z) return fail(GBW_ERR_INVALID, "source positions are NULL");
  if (pl.R > 0 && (!in->rcv_x || !in->rcv_y || !in->rcv_z))
    return fail(GBW_ERR_INVALID, "receiver positions are NULL");
  const bool rec = out->rec_p || out->rec_vx || out->rec_vy || out->rec_vz;
  if (rec && !(out->rec_p && out->rec_vx && out->rec_vy && out->rec_vz))
    return fail(GBW_ERR_INVALID, "rec_p, rec_vx, rec_vy and rec_vz must be given together");
  const Dims3& d = pl.d;
  // geometry, with the reference's IndexError semantics (vx has NX-1
  // columns, vy NY-1 rows, vz NZ-1 planes)
  std::vector<int> src(pl.S), rcell(std::max(1, pl.R), 0);
  for (int s = 0; s < pl.S; ++s) {
    const long long z = in->src_z[s] + d.nabc, y = in->src_y[s] + d.nabc, x = in->src_x[s] + d.nabc;
    if (z < 0 || z >= d.NZ || y < 0 || y >= d.NY || x < 0 || x >= d.NX)
      return fail(GBW_ERR_INDEX, "index out of range: source " + std::to_string(s));
    src[s] = (int)((z * d.NY + y) * d.PX + x);
  }
  for (int r = 0; r < pl.R; ++r) {
    const long long z = in->rcv_z[r] + d.nabc, y = in->rcv_y[r] + d.nabc, x = in->rcv_x[r] + d.nabc;
    if (z < 0 || z >= d.NZ - 1 || y < 0 || y >= d.NY - 1 || x < 0 || x >= d.NX - 1)
      return fail(GBW_ERR_INDEX, "index out of range: receiver " + std::to_string(r));
    rcell[r] = (int)((z * d.NY + y) * d.PX + x);
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_src), src.data(), src.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_rcell), rcell.data(), rcell.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  float* coef = at3<float>(workspace, pl.off_coef);
  coef3_kernel<<<(unsigned)((d.plane + 255) / 256), 256, 0, st>>>(d, in->vp, in->rho, in->damp,
                                                                    prob->dt, prob->dz, coef);
  LAUNCH_CHECK();
  const long long fstride = (long long)pl.S * d.plane;
  float* state = at3<float>(workspace, pl.off_state);
  if ((rc = fill3(state, 4 * fstride, 0.f, st))) return rc;
  const long long inner = (long long)d.nz * d.ny * d.nx;
  const bool energy = out->energy_p || out->energy_vx || out->energy_vy || out->energy_vz;
  float* E = at3<float>(workspace, pl.off_E);
  if (energy) {
    if ((rc = fill3(E, (long long)pl.groups * 4 * inner, 0.f, st))) return rc;
    float* eo[4] = {out->energy_p, out->energy_vx, out->energy_vy, out->energy_vz};
    for (float* e : eo)
      if (e && (rc = fill3(e, inner, 0.f, st))) return rc;
  }
  Step3Args a{};
  a.d = d; a.S = pl.S; a.spb = pl.spb; a.nt = pl.nt; a.dt = prob->dt;
  a.state = state; a.fstride = fstride; a.coef = coef;
  a.src_cell = at3<int>(workspace, pl.off_src);
  a.wav = in->src_v; a.wav_stride = prob->wavelet_shared ? 0 : prob->nt;
  a.energy = energy ? E : nullptr;
  const unsigned blocks = (unsigned)(((d.groups + T3 - 1) / T3 + 7) / 8 * 8);
  const dim3 grid(blocks, pl.groups);
  for (int it = 0; it < pl.nt; ++it) {
    a.it = it;
    {
      LaunchTimer timer("p3_step_kernel", st);
      p3_step_kernel<<<grid, T3, 0, st>>>(a);
    }
    LAUNCH_CHECK();
    {
      LaunchTimer timer("v3_step_kernel", st);
      v3_step_kernel<<<grid, T3, 0, st>>>(a);
    }
    LAUNCH_CHECK();
    if (rec && pl.R > 0) {
      record3_kernel<<<(unsigned)(((long long)pl.S * pl.R + 255) / 256), 256, 0, st>>>(
          d, pl.S, pl.R, pl.nt, it, state, fstride, at3<int>(workspace, pl.off_rcell), out->rec_p,
          out->rec_vx, out->rec_vy, out->rec_vz);
      LAUNCH_CHECK();
    }
    if (energy && (((it + 1) % pl.echunk == 0) || it + 1 == pl.nt)) {
      energy3_flush_kernel<<<(unsigned)((4 * inner + 255) / 256), 256, 0, st>>>(
          E, pl.groups, inner, out->energy_p, out->energy_vx, out->energy_vy, out->energy_vz);
      LAUNCH_CHECK();
    }
  }
  // the host vectors above were staged by hipMemcpyAsync before it returned;
  // wait anyway so they outlive the copies
  HIP_TRY(hipStreamSynchronize(st));
  return flush_timing(st);
}

}  // extern "C"
