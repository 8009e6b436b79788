// gbw_acoustic3d.hip -- MI355X (gfx950) kernels of the 3D acoustic propagator.
//
// Replaces the time stepping of GeoBrain's acoustic_forward_kernel_3d
// (geobrain/physics/wave/kernels/acoustic_3d.py:20-476): the 4th-order
// staggered-grid pressure / velocity scheme with split PML damping, point
// sources, receiver sampling and energy maps, and the hand-written adjoint
// that stands in for autograd through its checkpointed steps (vp, rho and
// wavelet gradients).
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
// path bit for bit.  The opt-in z-marching kernel (fwd3_march_kernel) fuses
// the p and v launches; the adjoint kernels follow the forward ones below.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <type_traits>
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

// Step kernels: XCD k (workgroups are dealt to the 8 XCDs round-robin) owns
// the y-slab [k NY / 8, (k + 1) NY / 8) of every plane and walks it in z, y,
// x order, 256 float4 groups per workgroup.  The workgroups resident on an
// XCD at one time then cover a few consecutive slab planes, so the planes
// above and below that the z-stencils read are fetched into its L2 by
// neighbouring workgroups at about the same time.  (Measured on a 240^3 grid
// with 8 shots: 6% faster than contiguous whole planes per XCD; cutting each
// slab into 2, 4 or 8 sub-slabs walked one after the other is slower.)
__device__ __forceinline__ bool lin_cell3(const Dims3& d, int& g, int& y, int& z, unsigned& q4) {
  const int k = (int)(blockIdx.x & 7);
  const int y0 = k * d.NY / 8, rows = (k + 1) * d.NY / 8 - y0;
  const long long l = (long long)(blockIdx.x >> 3) * T3 + threadIdx.x;
  if (rows <= 0 || l >= (long long)d.NZ * rows * d.GX) return false;
  g = (int)(l % d.GX);
  const long long t = l / d.GX;
  y = y0 + (int)(t % rows);
  z = (int)(t / rows);
  q4 = ((unsigned)z * (unsigned)d.NY + (unsigned)y) * (unsigned)d.GX + (unsigned)g;
  return true;
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
  const int* src_cell;     // [S * src_k] padded cell index in the plane of each point
  const float* wav; long long wav_stride;
  float* energy;           // [group][4][nz][ny][nx] or null
  float* dstore;           // recompute: [shot][plane] p-update bracket of this step, or null
  float* pstore;           // recompute: [shot][plane] p after this step's update, or null
  long long wav_kstride;   // wavelet row of point k of a shot: + k * wav_kstride
  int src_k;               // source points per shot (distinct cells)
};

// p on z in [f+1, NZ-3], y in [2, NY-3], x in [2, NX-3]; then the source.
// In-plane offsets are 32-bit (a plane of one field of one shot is < 2^28
// float4), so loads use a uniform base and a 32-bit lane offset.
// MP: several source points per shot (src_k > 1); the one-point instance
// compiles with src_k fixed at 1
template <bool MP>
__global__ __launch_bounds__(T3) void p3_step_kernel(Step3Args a) {
  const int K = MP ? a.src_k : 1;
  const Dims3 d = a.d;
  int g, y, z;
  unsigned q4;  // float4 group of the plane
  if (!lin_cell3(d, g, y, z, q4)) return;
  const int s_begin = blockIdx.y * a.spb, s_end = min(a.S, s_begin + a.spb);
  unsigned mask = 0;
  if (z >= d.f + 1 && z <= d.NZ - 3 && y >= 2 && y <= d.NY - 3)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 4 * g + e;
      if (x >= 2 && x <= d.NX - 3) mask |= 1u << e;
    }
  const float4* __restrict__ C = reinterpret_cast<const float4*>(a.coef);
  const float4 omk1 = ld4u(C, q4), a1 = ld4u(C + (d.plane >> 2), q4);
  const unsigned sy4 = (unsigned)d.GX, sz4 = (unsigned)d.NY * (unsigned)d.GX;
  const unsigned gl = g > 0 ? q4 - 1u : q4, gr = g + 1 < d.GX ? q4 + 1u : q4;
  const bool store = a.dstore != nullptr;
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    float4* __restrict__ P = reinterpret_cast<float4*>(a.state + so);
    const float4* __restrict__ VX = reinterpret_cast<const float4*>(a.state + a.fstride + so);
    const float4* __restrict__ VY = reinterpret_cast<const float4*>(a.state + 2 * a.fstride + so);
    const float4* __restrict__ VZ = reinterpret_cast<const float4*>(a.state + 3 * a.fstride + so);
    const int sc = a.src_cell[s * K];
    const bool src_here = (unsigned)(sc >> 2) == q4;
    // further points of the shot (acoustic_3d.py:171-176) in this group
    bool more_here = false;
    if constexpr (MP)
      for (int k = 1; k < K; ++k)
        more_here = more_here || (unsigned)(a.src_cell[s * K + k] >> 2) == q4;
    if (!mask && !src_here && !more_here && !store) continue;
    float4 p = ld4u(P, q4);
    float4 dsum = make_float4(0.f, 0.f, 0.f, 0.f);
    if (mask) {
      const float4 xa = ld4u(VX, gl), xb = ld4u(VX, q4), xc = ld4u(VX, gr);
      const float vxs[12] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w};
      const float4 ym2 = ld4u(VY, q4 - 2u * sy4), ym1 = ld4u(VY, q4 - sy4), y00 = ld4u(VY, q4),
                   yp1 = ld4u(VY, q4 + sy4);
      const float4 zm2 = ld4u(VZ, q4 - 2u * sz4), zm1 = ld4u(VZ, q4 - sz4), z00 = ld4u(VZ, q4),
                   zp1 = ld4u(VZ, q4 + sz4);
      float4 pn;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float dvx = d4(vxs[4 + e], vxs[3 + e], vxs[5 + e], vxs[2 + e]);
        const float dvy = d4(g3(y00, e), g3(ym1, e), g3(yp1, e), g3(ym2, e));
        const float dvz = d4(g3(z00, e), g3(zm1, e), g3(zp1, e), g3(zm2, e));
        float sum = dvx + dvy;
        sum = sum + dvz;
        s3(dsum, e, sum);
        const float u = g3(omk1, e) * g3(p, e);
        s3(pn, e, u - g3(a1, e) * sum);
      }
      if (mask == 15u) {
        p = pn;
      } else {  // cells outside the region keep their value (and a zero bracket)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool in = (mask >> e) & 1u;
          s3(p, e, in ? g3(pn, e) : g3(p, e));
          s3(dsum, e, in ? g3(dsum, e) : 0.f);
        }
      }
    }
    if (src_here) {
      const int e = sc & 3;
      s3(p, e, g3(p, e) + a.dt * a.wav[(long long)s * a.wav_stride + a.it]);
    }
    if (more_here)  // distinct cells: each adds to its own
      for (int k = 1; k < K; ++k) {
        const int sk = a.src_cell[s * K + k];
        if ((unsigned)(sk >> 2) != q4) continue;
        const int e = sk & 3;
        s3(p, e, g3(p, e) + a.dt * a.wav[(long long)s * a.wav_stride + k * a.wav_kstride + a.it]);
      }
    st4u(P, q4, p);
    if (store) {
      st4u(reinterpret_cast<float4*>(a.dstore + so), q4, dsum);
      if (a.pstore) st4u(reinterpret_cast<float4*>(a.pstore + so), q4, p);
    }
  }
}

// the p mirror, vx / vy / vz on their regions, the vz mirror, energy sums
__global__ __launch_bounds__(T3) void v3_step_kernel(Step3Args a) {
  const Dims3 d = a.d;
  int g, y, z;
  unsigned q4;
  if (!lin_cell3(d, g, y, z, q4)) return;
  const int s_begin = blockIdx.y * a.spb, s_end = min(a.S, s_begin + a.spb);
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
  const float4 omk2 = ld4u(C + 2 * p4, q4), omk3 = ld4u(C + 3 * p4, q4);
  const float4 omk4 = ld4u(C + 4 * p4, q4), a2 = ld4u(C + 5 * p4, q4);
  const unsigned sy4 = (unsigned)d.GX, sz4 = (unsigned)d.NY * (unsigned)d.GX;
  const unsigned gl = g > 0 ? q4 - 1u : q4, gr = g + 1 < d.GX ? q4 + 1u : q4;
  float4 ep = make_float4(0.f, 0.f, 0.f, 0.f), ex = ep, ey = ep, ez = ep;
  auto add_sq = [](float4& acc, const float4& v) {
    acc = make_float4(acc.x + v.x * v.x, acc.y + v.y * v.y, acc.z + v.z * v.z, acc.w + v.w * v.w);
  };
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    float4* __restrict__ P = reinterpret_cast<float4*>(a.state + so);
    float4* __restrict__ VX = reinterpret_cast<float4*>(a.state + a.fstride + so);
    float4* __restrict__ VY = reinterpret_cast<float4*>(a.state + 2 * a.fstride + so);
    float4* __restrict__ VZ = reinterpret_cast<float4*>(a.state + 3 * a.fstride + so);
    if (mirror_p) {  // no velocity of row f-1 is updated; nothing else reads this row here
      const float4 v = ld4u(P, q4 + 2u * sz4);
      st4u(P, q4, make_float4(-v.x, -v.y, -v.z, -v.w));
      continue;
    }
    const float4 pb = ld4u(P, q4);
    float4 vx = ld4u(VX, q4), vy = ld4u(VY, q4), vz = ld4u(VZ, q4);
    if (mx) {
      const float4 pa = ld4u(P, gl), pc = ld4u(P, gr);
      const float ps[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w, pc.x, pc.y, pc.z, pc.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float u = g3(omk2, e) * g3(vx, e);
        const float v = u - g3(a2, e) * d4(ps[5 + e], ps[4 + e], ps[6 + e], ps[3 + e]);
        s3(vx, e, ((mx >> e) & 1u) ? v : g3(vx, e));
      }
      st4u(VX, q4, vx);
    }
    if (my) {
      const float4 pu = ld4u(P, q4 - sy4), pd1 = ld4u(P, q4 + sy4), pd2 = ld4u(P, q4 + 2u * sy4);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float u = g3(omk3, e) * g3(vy, e);
        const float v = u - g3(a2, e) * d4(g3(pd1, e), g3(pb, e), g3(pd2, e), g3(pu, e));
        s3(vy, e, ((my >> e) & 1u) ? v : g3(vy, e));
      }
      st4u(VY, q4, vy);
    }
    if (mz) {
      // at z = f the plane above is the mirrored row f-1: read it as -p[f+1]
      const float4 pd1 = ld4u(P, q4 + sz4), pd2 = ld4u(P, q4 + 2u * sz4);
      const float4 pu = vz_up ? make_float4(-pd1.x, -pd1.y, -pd1.z, -pd1.w) : ld4u(P, q4 - sz4);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float u = g3(omk4, e) * g3(vz, e);
        const float v = u - g3(a2, e) * d4(g3(pd1, e), g3(pb, e), g3(pd2, e), g3(pu, e));
        s3(vz, e, ((mz >> e) & 1u) ? v : g3(vz, e));
      }
      st4u(VZ, q4, vz);
    }
    if (vz_up) st4u(VZ, q4 - sz4, vz);  // the whole of row f, updated or not
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
    // nabc and nx multiples of 4: the group lies wholly inside the interior
    // (me == 15) and its four cells are one aligned float4 of each map
    if (((d.nabc | d.nx) & 3) == 0) {
      float4* e4 = reinterpret_cast<float4*>(eg + ob + (4 * g - d.nabc));
      const long long i4 = inner >> 2;
      const float4 acc[4] = {ep, ex, ey, ez};
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const float4 o = e4[m * i4];
        e4[m * i4] = make_float4(o.x + acc[m].x, o.y + acc[m].y, o.z + acc[m].z, o.w + acc[m].w);
      }
      return;
    }
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

// ---------------------------------------------------------------------------
// Adjoint (reverse) step.  The transpose of step `it` runs in three parts:
//   q part   the adjoint sources of step `it`, vz[f-1] = vz[f] and the vz / vy / vx
//            updates: q = alpha2 lam_v, d alpha2 -= lam_v bracket(p'),
//            lam_v *= 1 - kappa
//   p part   lam_p += -D^T q, p[f-1] = -p[f+1], the source term (wavelet
//            gradient), the p update: r = alpha1 lam_p, d alpha1 -= lam_p D,
//            lam_p *= 1 - kappa1
//   v part   lam_v += -B^T r
// Two launches per reverse step: adj3_vq_kernel runs the v part of step it+1
// and the q part of step `it` in one pass over lam_v (both are pointwise in
// lam_v once r of step it+1 is written), adj3_p_kernel the p part.  lam_v is
// stored before the q part's scaling: adj3_p_kernel forms q = alpha2 lam_v on
// its stencil as it reads it, and the next adj3_vq_kernel applies the scaling,
// so q is never stored.  The v part of step 0 changes no gradient and is not run.
// ---------------------------------------------------------------------------

struct Adj3Args {
  Dims3 d;
  int S, spb, it, nt;
  float dt;
  int vpart;               // adj3_vq_kernel: run the v part of step it+1 first
  float* lam;              // [field p|vx|vy|vz][shot][plane]
  long long fstride;       // S * plane
  float* r;                // [shot][plane]
  const float* coef;
  const float* dstep;      // p-update bracket of step `it`, [shot][plane]
  const float* pstep;      // p after step `it`, [shot][plane]; null without the density gradient
  float* gA1;              // [group][plane]
  float* gA2;              // [group][plane] or null
  const int* src_cell;
  float* gw;               // grad_src_v (S, nt), or null
  float* gws;              // [S] source terms of this step for a shared wavelet, or null
  // adjoint sources: receiver cells sorted by cell, the receivers of each in
  // receiver order, and per (z, y) row the range of its cells; null: none
  const int* rowstart;     // [NZ * NY + 1] into ucell
  const int* ucell;        // in-plane cell index
  const int* ustart;       // [ncell + 1] into urcv
  const int* urcv;         // receiver numbers
  int R;
  const float* g_p; const float* g_vx; const float* g_vy; const float* g_vz;  // (S, nt, R) or null
  int src_k;               // source points per shot (gw row s * src_k + k)
};

// float4 array addressed by a 32-bit element index (byte offset i * 16 < 2^32):
// loads and stores take a uniform base and a 32-bit lane offset
template <typename T>
struct A4 {
  T* p;
  __device__ __forceinline__ T& operator[](unsigned i) const {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return *reinterpret_cast<T*>(reinterpret_cast<B*>(p) + (i << 4));
  }
};

__device__ __forceinline__ float4 ld4(const A4<const float4>& a, unsigned i, bool ok) {
  return ok ? a[i] : make_float4(0.f, 0.f, 0.f, 0.f);
}

struct Cell3 {
  int g, y, z;
  unsigned q4;
};

__device__ __forceinline__ bool cell3(const Dims3& d, Cell3& c) {
  return lin_cell3(d, c.g, c.y, c.z, c.q4);
}

__device__ __forceinline__ float4 ld4(const float4* p, long long i, bool ok) {
  return ok ? p[i] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// receiver cells of the float4 group q4 (plane row `row`): ucell[u0 .. u1)
__device__ __forceinline__ void rcv_cells(const Adj3Args& a, int row, unsigned q4, int& u0,
                                          int& u1) {
  int lo = a.rowstart[row], hi = a.rowstart[row + 1];
  const int end = hi, c0 = (int)(q4 << 2);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a.ucell[mid] < c0) lo = mid + 1;
    else hi = mid;
  }
  u0 = u1 = lo;
  while (u1 < end && a.ucell[u1] < c0 + 4) ++u1;
}

// v += the adjoint sources of cells ucell[u0 .. u1) (group starting at cell c0),
// the receivers of a cell in receiver order
__device__ __forceinline__ void add_sources(const Adj3Args& a, const float* gf, long long base,
                                            int u0, int u1, int c0, float4& v) {
  for (int u = u0; u < u1; ++u) {
    const int e = a.ucell[u] - c0;
    float x = g3(v, e);
    for (int j = a.ustart[u]; j < a.ustart[u + 1]; ++j) x = x + gf[base + a.urcv[j]];
    s3(v, e, x);
  }
}

// v part of step it+1, adjoint sources of step `it`, q part of step `it`
__global__ __launch_bounds__(T3) void adj3_vq_kernel(Adj3Args a) {
  const Dims3 d = a.d;
  Cell3 c;
  if (!cell3(d, c)) return;
  const int g = c.g, y = c.y, z = c.z;
  const unsigned q4 = (unsigned)c.q4;
  const unsigned sy4 = (unsigned)d.GX, sz4 = (unsigned)d.NY * (unsigned)d.GX;
  const bool zx = z >= d.f && z <= d.NZ - 2, zz = z >= d.f && z <= d.NZ - 3;
  unsigned mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int x = 4 * g + e;
    if (zx && y >= 1 && y <= d.NY - 2 && x >= 1 && x <= d.NX - 3) mx |= 1u << e;
    if (zx && y >= 1 && y <= d.NY - 3 && x >= 1 && x <= d.NX - 2) my |= 1u << e;
    if (zz && y >= 1 && y <= d.NY - 2 && x >= 1 && x <= d.NX - 2) mz |= 1u << e;
  }
  const bool fold = d.fs && z == d.f;       // transpose of vz[f-1] = vz[f]
  // with a free surface lam_vz of plane f-1 is folded into plane f by that
  // plane's thread, which recomputes it; here it stays 0 in memory
  const bool above = d.fs && z == d.f - 1;
  // the v part reaches z in [f-1, NZ-2], y <= NY-2 (r is zero beyond)
  const bool vin = a.vpart && z >= d.f - 1 && z <= d.NZ - 2 && y <= d.NY - 2;
  const bool vin_up = fold && a.vpart && y <= d.NY - 2;  // the v part of plane f-1's cell
  int u0 = 0, u1 = 0, w0 = 0, w1 = 0;  // receiver cells of this group and of plane f-1's
  if (a.rowstart) {
    rcv_cells(a, z * d.NY + y, q4, u0, u1);
    if (fold) rcv_cells(a, (z - 1) * d.NY + y, q4 - sz4, w0, w1);
  }
  const bool rcv = u1 > u0;
  if (!(mx | my | mz) && !fold && !vin && !rcv) return;
  const float4* __restrict__ C = reinterpret_cast<const float4*>(a.coef);
  const long long p4 = d.plane >> 2;
  const float4 omk2 = ld4u(C + 2 * p4, q4), omk3 = ld4u(C + 3 * p4, q4);
  const float4 omk4 = ld4u(C + 4 * p4, q4), a2 = ld4u(C + 5 * p4, q4);
  const unsigned gl = g > 0 ? q4 - 1u : q4, gr = g + 1 < d.GX ? q4 + 1u : q4;
  const bool rho = a.gA2 != nullptr;
  const int s_begin = blockIdx.y * a.spb, s_end = min(a.S, s_begin + a.spb);
  const int c0 = (int)(q4 << 2);
  float4 acc2 = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const A4<float4> LX{reinterpret_cast<float4*>(a.lam + a.fstride + so)};
    const A4<float4> LY{reinterpret_cast<float4*>(a.lam + 2 * a.fstride + so)};
    const A4<float4> LZ{reinterpret_cast<float4*>(a.lam + 3 * a.fstride + so)};
    const A4<const float4> Rr{reinterpret_cast<const float4*>(a.r + so)};
    // lam_v as the q part of step it+1 left it, before its scaling
    // lam_v *= 1 - kappa on the v regions, which is applied here
    float4 lx = LX[q4], ly = LY[q4], lz = above ? make_float4(0.f, 0.f, 0.f, 0.f) : LZ[q4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if ((mz >> e) & 1u) s3(lz, e, g3(omk4, e) * g3(lz, e));
      if ((my >> e) & 1u) s3(ly, e, g3(omk3, e) * g3(ly, e));
      if ((mx >> e) & 1u) s3(lx, e, g3(omk2, e) * g3(lx, e));
    }
    if (vin) {  // lam_v += c1 (r[i+1] - r[i]) + c2 (r[i+2] - r[i-1]) along each axis
      const float4 r0 = Rr[q4];
      const float4 xa = ld4(Rr, q4 - 1, g >= 1), xc = ld4(Rr, q4 + 1, g + 1 < d.GX);
      const float rs[12] = {xa.x, xa.y, xa.z, xa.w, r0.x, r0.y, r0.z, r0.w, xc.x, xc.y, xc.z, xc.w};
      const float4 ym1 = ld4(Rr, q4 - sy4, y >= 1), yp1 = Rr[q4 + sy4];
      const float4 yp2 = ld4(Rr, q4 + 2 * sy4, y + 2 < d.NY);
      const float4 zm1 = ld4(Rr, q4 - sz4, z >= 1), zp1 = Rr[q4 + sz4];
      const float4 zp2 = ld4(Rr, q4 + 2 * sz4, z + 2 < d.NZ);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        s3(lx, e, g3(lx, e) + d4(rs[5 + e], rs[4 + e], rs[6 + e], rs[3 + e]));
        s3(ly, e, g3(ly, e) + d4(g3(yp1, e), g3(r0, e), g3(yp2, e), g3(ym1, e)));
        s3(lz, e, g3(lz, e) + d4(g3(zp1, e), g3(r0, e), g3(zp2, e), g3(zm1, e)));
      }
    }
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    const long long base = ((long long)s * a.nt + a.it) * a.R;
    if (rcv) {
      if (a.g_p) {
        const A4<float4> LP{reinterpret_cast<float4*>(a.lam + so)};
        float4 lp = LP[q4];
        add_sources(a, a.g_p, base, u0, u1, c0, lp);
        LP[q4] = lp;
      }
      if (a.g_vx) add_sources(a, a.g_vx, base, u0, u1, c0, lx);
      if (a.g_vy) add_sources(a, a.g_vy, base, u0, u1, c0, ly);
      if (a.g_vz && !above) add_sources(a, a.g_vz, base, u0, u1, c0, lz);
    }
    if (fold) {
      // lam_vz of plane f-1 after the v part and its sources (it starts the step at 0)
      float4 up = zero;
      if (vin_up) {
        const unsigned qm = q4 - sz4;
        const int zm = z - 1;
        const float4 r0 = Rr[qm], zm1 = ld4(Rr, qm - sz4, zm >= 1), zp1 = Rr[qm + sz4];
        const float4 zp2 = ld4(Rr, qm + 2 * sz4, zm + 2 < d.NZ);
#pragma unroll
        for (int e = 0; e < 4; ++e) s3(up, e, 0.f + d4(g3(zp1, e), g3(r0, e), g3(zp2, e), g3(zm1, e)));
      }
      if (w1 > w0 && a.g_vz) add_sources(a, a.g_vz, base, w0, w1, c0 - (int)(sz4 << 2), up);
      lz = make_float4(lz.x + up.x, lz.y + up.y, lz.z + up.z, lz.w + up.w);
    }
    if (rho) {  // d alpha2 -= lam_v bracket(p'), the brackets of the v updates from p after the step
      float pxs[12];
      float4 pb = zero, pyu = zero, pyd1 = zero, pyd2 = zero, pzu = zero, pzd1 = zero, pzd2 = zero;
      const A4<const float4> P1{reinterpret_cast<const float4*>(a.pstep + so)};
      pb = P1[q4];
      if (mx) {
        const float4 pa = P1[gl], pc = P1[gr];
        pxs[0] = pa.x; pxs[1] = pa.y; pxs[2] = pa.z; pxs[3] = pa.w;
        pxs[8] = pc.x; pxs[9] = pc.y; pxs[10] = pc.z; pxs[11] = pc.w;
      }
      pxs[4] = pb.x; pxs[5] = pb.y; pxs[6] = pb.z; pxs[7] = pb.w;
      if (my) {
        pyu = P1[q4 - sy4];
        pyd1 = P1[q4 + sy4];
        pyd2 = P1[q4 + 2 * sy4];
      }
      if (mz) {
        pzd1 = P1[q4 + sz4];
        pzd2 = P1[q4 + 2 * sz4];
        // the forward read the mirrored plane f-1 as -p[f+1]
        pzu = fold ? make_float4(-pzd1.x, -pzd1.y, -pzd1.z, -pzd1.w) : P1[q4 - sz4];
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float acc = g3(acc2, e);
        if ((mz >> e) & 1u) acc = acc - g3(lz, e) * d4(g3(pzd1, e), g3(pb, e), g3(pzd2, e), g3(pzu, e));
        if ((my >> e) & 1u) acc = acc - g3(ly, e) * d4(g3(pyd1, e), g3(pb, e), g3(pyd2, e), g3(pyu, e));
        if ((mx >> e) & 1u) acc = acc - g3(lx, e) * d4(pxs[5 + e], pxs[4 + e], pxs[6 + e], pxs[3 + e]);
        s3(acc2, e, acc);
      }
    }
    // stored before the q part's scaling: adj3_p_kernel forms q = alpha2 lam_v
    // from it, and the next launch scales it
    const bool upd = vin || rcv;
    if (mx || upd) LX[q4] = lx;
    if (my || upd) LY[q4] = ly;
    if ((mz || fold || upd) && !above) LZ[q4] = lz;
  }
  if (rho && (mx | my | mz)) {
    const A4<float4> G{reinterpret_cast<float4*>(a.gA2 + (long long)blockIdx.y * d.plane)};
    const float4 o = G[q4];
    G[q4] = make_float4(o.x + acc2.x, o.y + acc2.y, o.z + acc2.z, o.w + acc2.w);
  }
}

// The update-region masks of the q = alpha2 lam_v that the lam_p gather of one
// cell group reads, the same for every shot (q is zero off each velocity's
// region and outside the grid)
struct QStencil {
  unsigned mz[4], my[4], mx[3];  // element masks: planes z-2 .. z+1; rows y-2 .. y+1; groups g-1 .. g+1
};

__device__ __forceinline__ QStencil q_stencil(const Dims3& d, int z, int y, int g) {
  QStencil t;
  unsigned xyz = 0;  // x in [1, NX-2]: the vy and vz regions
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (4 * g + e >= 1 && 4 * g + e <= d.NX - 2) xyz |= 1u << e;
  const bool yin = y >= 1 && y <= d.NY - 2, zx = z >= d.f && z <= d.NZ - 2;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int zz = z - 2 + i, yy = y - 2 + i;
    t.mz[i] = (zz >= d.f && zz <= d.NZ - 3 && yin) ? xyz : 0u;
    t.my[i] = (zx && yy >= 1 && yy <= d.NY - 3) ? xyz : 0u;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int h = g - 1 + i;
    unsigned m = 0;
    if (zx && yin && h >= 0 && h < d.GX)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (4 * h + e >= 1 && 4 * h + e <= d.NX - 3) m |= 1u << e;
    t.mx[i] = m;
  }
  return t;
}

// q = alpha2 lam_v on the elements of m, else 0.  Every q of the gather is
// loaded unconditionally (from the group itself where m is empty, so no
// address leaves the volume) and masked afterwards: loads under a branch
// would each wait for their own round trip.  alpha2 is re-read for every
// shot, which measured faster than holding the gather's nine alpha2 groups
// in registers across the shot loop.
__device__ __forceinline__ float4 q_of(const A4<const float4>& A2, const A4<const float4>& L,
                                       unsigned i, unsigned m, unsigned own) {
  const unsigned j = m ? i : own;
  const float4 c = A2[j], l = L[j];
  float4 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) s3(o, e, ((m >> e) & 1u) ? g3(c, e) * g3(l, e) : 0.f);
  return o;
}

// -D^T q at the cell group: c1 (q[i] - q[i-1]) + c2 (q[i+1] - q[i-2]) per axis,
// added z, then y, then x; lam_v is read as adj3_vq_kernel left it.  Also
// returns q_z of plane z-1 (the lam_p gather of plane z-2 needs it).
__device__ __forceinline__ float4 gather_q(const Dims3& d, const QStencil& t,
                                           const A4<const float4>& A2,
                                           const A4<const float4>& LX,
                                           const A4<const float4>& LY,
                                           const A4<const float4>& LZ, unsigned q4,
                                           float4& qz_up) {
  const unsigned sy4 = (unsigned)d.GX, sz4 = (unsigned)d.NY * (unsigned)d.GX;
  const float4 zm2 = q_of(A2, LZ, q4 - 2 * sz4, t.mz[0], q4);
  const float4 zm1 = q_of(A2, LZ, q4 - sz4, t.mz[1], q4);
  const float4 z0 = q_of(A2, LZ, q4, t.mz[2], q4);
  const float4 zp1 = q_of(A2, LZ, q4 + sz4, t.mz[3], q4);
  const float4 ym2 = q_of(A2, LY, q4 - 2 * sy4, t.my[0], q4);
  const float4 ym1 = q_of(A2, LY, q4 - sy4, t.my[1], q4);
  const float4 y0 = q_of(A2, LY, q4, t.my[2], q4);
  const float4 yp1 = q_of(A2, LY, q4 + sy4, t.my[3], q4);
  const float4 xa = q_of(A2, LX, q4 - 1, t.mx[0], q4);
  const float4 xb = q_of(A2, LX, q4, t.mx[1], q4);
  const float4 xc = q_of(A2, LX, q4 + 1, t.mx[2], q4);
  const float xs[12] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w, xc.x, xc.y, xc.z, xc.w};
  float4 out;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float acc = 0.f;
    acc = acc + d4(g3(z0, e), g3(zm1, e), g3(zp1, e), g3(zm2, e));
    acc = acc + d4(g3(y0, e), g3(ym1, e), g3(yp1, e), g3(ym2, e));
    acc = acc + d4(xs[4 + e], xs[3 + e], xs[5 + e], xs[2 + e]);
    s3(out, e, acc);
  }
  qz_up = zm1;
  return out;
}

// MP: several source points per shot (src_k > 1); the one-point instance
// compiles with src_k fixed at 1
template <bool MP>
__global__ __launch_bounds__(T3) void adj3_p_kernel(Adj3Args a) {
  const int K = MP ? a.src_k : 1;
  const Dims3 d = a.d;
  Cell3 c;
  if (!cell3(d, c)) return;
  const int g = c.g, y = c.y, z = c.z;
  const unsigned q4 = (unsigned)c.q4;
  unsigned mp = 0;
  if (z >= d.f + 1 && z <= d.NZ - 3 && y >= 2 && y <= d.NY - 3)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 4 * g + e;
      if (x >= 2 && x <= d.NX - 3) mp |= 1u << e;
    }
  const bool fold = d.fs && z == d.f + 1;   // transpose of p[f-1] = -p[f+1]
  const bool mirrored = d.fs && z == d.f - 1;  // zeroed by the fold at f+1
  const float4* __restrict__ C = reinterpret_cast<const float4*>(a.coef);
  const long long p4 = d.plane >> 2;
  const float4 omk1 = ld4u(C, q4), a1 = ld4u(C + p4, q4);
  const unsigned sz4 = (unsigned)d.NY * (unsigned)d.GX;
  const A4<const float4> A2{C + 5 * p4};
  const QStencil qs = q_stencil(d, z, y, g);
  const int s_begin = blockIdx.y * a.spb, s_end = min(a.S, s_begin + a.spb);
  float4 acc1 = make_float4(0.f, 0.f, 0.f, 0.f);
  // source cells one shot ahead: a load at the top of the shot would be
  // waited for before anything else of the shot is issued
  int sc_next = a.src_cell[s_begin * K];
  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const int sc = sc_next;
    {
      // a per-lane index: a uniform one makes the value scalar, read (and
      // so waited for) as soon as it is loaded
      int sn = min(s + 1, s_end - 1);
      asm volatile("" : "+v"(sn));
      sc_next = a.src_cell[sn * K];
    }
    const bool src_here = (unsigned)(sc >> 2) == q4;
    if (mirrored) {  // the source term on the mirrored plane is overwritten: no gradient
      if (src_here) {
        if (a.gws) a.gws[s] = 0.f;
        if (a.gw) a.gw[(long long)s * K * a.nt + a.it] = 0.f;
      }
      continue;  // further points here keep their zeroed samples
    }
    const A4<float4> LP{reinterpret_cast<float4*>(a.lam + so)};
    const A4<const float4> LX{reinterpret_cast<const float4*>(a.lam + a.fstride + so)};
    const A4<const float4> LY{reinterpret_cast<const float4*>(a.lam + 2 * a.fstride + so)};
    const A4<const float4> LZ{reinterpret_cast<const float4*>(a.lam + 3 * a.fstride + so)};
    float4 lp = LP[q4];
    float4 qf;  // q_z of plane z-1
    const float4 gq = gather_q(d, qs, A2, LX, LY, LZ, q4, qf);
    lp = make_float4(lp.x + gq.x, lp.y + gq.y, lp.z + gq.z, lp.w + gq.w);
    if (fold) {
      float4 lm = LP[q4 - 2 * sz4];
      // the gather at plane f-1 (= z-2): of the q it reads only q_z of plane f
      // (= z-1) lies on a velocity region; the zeros are added as the gather does
      float4 gm;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float acc = 0.f;
        acc = acc + d4(0.f, 0.f, g3(qf, e), 0.f);
        acc = acc + d4(0.f, 0.f, 0.f, 0.f);
        acc = acc + d4(0.f, 0.f, 0.f, 0.f);
        s3(gm, e, acc);
      }
      lm = make_float4(lm.x + gm.x, lm.y + gm.y, lm.z + gm.z, lm.w + gm.w);
      lp = make_float4(lp.x - lm.x, lp.y - lm.y, lp.z - lm.z, lp.w - lm.w);
      LP[q4 - 2 * sz4] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (src_here) {
      const float v = a.dt * g3(lp, sc & 3);
      if (a.gws) a.gws[s] = v;
      if (a.gw) a.gw[(long long)s * K * a.nt + a.it] = v;
    }
    if constexpr (MP) {  // further points: a row of samples each
      if (a.gw)
        for (int k = 1; k < K; ++k) {
          const int sk = a.src_cell[s * K + k];
          if ((unsigned)(sk >> 2) == q4)
            a.gw[((long long)s * K + k) * a.nt + a.it] = a.dt * g3(lp, sk & 3);
        }
    }
    if (mp) {
      const float4 D = ld4u(reinterpret_cast<const float4*>(a.dstep + so), q4);
      float4 rr = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((mp >> e) & 1u)) continue;
        const float l = g3(lp, e);
        s3(rr, e, g3(a1, e) * l);
        s3(acc1, e, g3(acc1, e) - l * g3(D, e));
        s3(lp, e, g3(omk1, e) * l);
      }
      st4u(reinterpret_cast<float4*>(a.r + so), q4, rr);
    }
    LP[q4] = lp;
  }
  if (mp && a.gA1) {
    const A4<float4> G{reinterpret_cast<float4*>(a.gA1 + (long long)blockIdx.y * d.plane)};
    const float4 o = G[q4];
    G[q4] = make_float4(o.x + acc1.x, o.y + acc1.y, o.z + acc1.z, o.w + acc1.w);
  }
}

// ---------------------------------------------------------------------------
// Fused z-marching step (opt-in, step_kernel 1).  A workgroup owns a 64 x 16 (x, y)
// tile, a range of z planes and a group of SPB shots, and marches down z.  At
// plane zp it computes p' on the tile and on the halo cells the velocity
// stencils read (one column left, two right, one row above, two below) into
// LDS, then vx', vy' at zp and vz' at zp - 2, whose stencil needs p' of
// zp - 3 .. zp (a register queue).  vz of zp - 2 .. zp + 1, the z reach of the
// p stencil, is a register queue too, so every field is read from HBM once per
// step.  The state is out of place (read A, write B): halo cells are
// recomputed from A while neighbouring tiles write B.
// ---------------------------------------------------------------------------
constexpr int MT3 = 256;           // threads: 16 float4 groups x 16 rows
constexpr int MLR = 19, MLG = 18;  // LDS tile: rows y0-1 .. y0+17, groups g0-1 .. g0+16

struct March3Args {
  Dims3 d;
  int S, it, nt;
  float dt;
  const float* A;          // state before the step, [field][shot][plane]
  float* B;                // state after the step
  long long fstride;       // S * plane
  const float* coef;
  const int* src_cell;
  const float* wav; long long wav_stride;
  float* energy;           // [group][4][nz][ny][nx] or null
  float* dstore;           // recompute: p-update bracket [shot][plane], or null
  float* pstore;           // recompute: p after the update [shot][plane], or null
  int tiles_x, tiles_y, zc, groups;
  long long units;
};

__device__ __forceinline__ float4 z4() { return make_float4(0.f, 0.f, 0.f, 0.f); }

// p' of the float4 group (z, y, g) of one shot: the p-region update from the
// old state, then the source term.  vz holds vz of planes z-2 .. z+1; the
// group's own vx and vy are returned for the velocity update.
template <bool DSUM>
__device__ __forceinline__ float4 march_p(const Dims3& d, const float4* P, const float4* VX,
                                          const float4* VY, const float4* vz, int z, int y,
                                          int g, float4 omk1, float4 a1, int sc, float w,
                                          float dt, float4& vxc, float4& vyc, float4& dsum) {
  const long long sy4 = d.GX;
  const long long q4 = ((long long)z * d.NY + y) * d.GX + g;
  float4 p = P[q4];
  vxc = VX[q4];
  vyc = VY[q4];
  unsigned mask = 0;
  if (z >= d.f + 1 && z <= d.NZ - 3 && y >= 2 && y <= d.NY - 3)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 4 * g + e;
      if (x >= 2 && x <= d.NX - 3) mask |= 1u << e;
    }
  if (mask) {
    const float4 xa = ld4(VX, q4 - 1, g >= 1), xc = ld4(VX, q4 + 1, g + 1 < d.GX);
    const float vxs[12] = {xa.x, xa.y, xa.z, xa.w, vxc.x, vxc.y, vxc.z, vxc.w,
                           xc.x, xc.y, xc.z, xc.w};
    const float4 ym2 = VY[q4 - 2 * sy4], ym1 = VY[q4 - sy4], yp1 = VY[q4 + sy4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (!((mask >> e) & 1u)) continue;
      const float dvx = d4(vxs[4 + e], vxs[3 + e], vxs[5 + e], vxs[2 + e]);
      const float dvy = d4(g3(vyc, e), g3(ym1, e), g3(yp1, e), g3(ym2, e));
      const float dvz = d4(g3(vz[2], e), g3(vz[1], e), g3(vz[3], e), g3(vz[0], e));
      float sum = dvx + dvy;
      sum = sum + dvz;
      if (DSUM) s3(dsum, e, sum);
      const float u = g3(omk1, e) * g3(p, e);
      s3(p, e, u - g3(a1, e) * sum);
    }
  }
  if (sc >= 0 && (long long)(sc >> 2) == q4) {
    const int e = sc & 3;
    s3(p, e, g3(p, e) + dt * w);
  }
  return p;
}

// energy partial += acc over the interior lanes of group (z, y, g)
__device__ __forceinline__ void energy_add(const Dims3& d, float* eg, int z, int y, int g,
                                           float4 acc) {
  if (z < d.nabc || z >= d.nabc + d.nz || y < d.nabc || y >= d.nabc + d.ny) return;
  const long long ob = ((long long)(z - d.nabc) * d.ny + (y - d.nabc)) * d.nx - d.nabc;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int x = 4 * g + e;
    if (x < d.nabc || x >= d.nabc + d.nx) continue;
    eg[ob + x] = eg[ob + x] + g3(acc, e);
  }
}

template <int SPB, bool REC>
__global__ __launch_bounds__(MT3) void fwd3_march_kernel(March3Args a) {
  __shared__ float4 lp[SPB][2][MLR][MLG];
  const Dims3 d = a.d;
  // XCD-aware: consecutive units (the shot groups of a tile, then neighbouring tiles) share an XCD
  const long long L = (long long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (L >= a.units) return;
  const int grp = (int)(L % a.groups);
  long long t = L / a.groups;
  const int tx = (int)(t % a.tiles_x);
  t /= a.tiles_x;
  const int tyi = (int)(t % a.tiles_y);
  const int zci = (int)(t / a.tiles_y);
  const int g0 = tx * 16, y0 = tyi * 16;
  const int zb = zci * a.zc, ze = min(d.NZ, zb + a.zc);
  const int s0 = grp * SPB, ns = min(SPB, a.S - s0);
  const int tid = threadIdx.x, gx = tid & 15, ty = tid >> 4;
  const int g = g0 + gx, y = y0 + ty;
  const bool tin = y < d.NY;
  // halo item of threads 0..79: left column, right column, the row above, the two rows below
  int hg = 0, hy = 0;
  if (tid < 16) { hg = g0 - 1; hy = y0 + tid; }
  else if (tid < 32) { hg = g0 + 16; hy = y0 + tid - 16; }
  else if (tid < 48) { hg = g0 + tid - 32; hy = y0 - 1; }
  else if (tid < 80) { hg = g0 + (tid - 48) % 16; hy = y0 + 16 + (tid - 48) / 16; }
  const bool hslot = tid < 80;
  const bool hin = hslot && hg >= 0 && hg < d.GX && hy >= 0 && hy < d.NY;
  const int hr = hy - y0 + 1, hc = hg - g0 + 1;
  const long long sz4 = (long long)d.NY * d.GX, p4 = d.plane >> 2;
  const long long f4 = a.fstride >> 2;
  const float4* __restrict__ C = reinterpret_cast<const float4*>(a.coef);
  const float4* __restrict__ A4 = reinterpret_cast<const float4*>(a.A);
  float4* __restrict__ B4 = reinterpret_cast<float4*>(a.B);
  // this chunk holds the mirrored plane f-1: it also writes p[f-1] = -p'[f+1], vz[f-1] = vz'[f]
  const bool fsm = d.fs && d.f - 1 >= zb && d.f - 1 < ze;
  const int zs = max(zb - 1, 0);
  const int zend = fsm ? max(ze + 1, d.f + 2) : ze + 1;
  int sc[SPB];
  float wv[SPB];
  long long so4[SPB];
#pragma unroll
  for (int k = 0; k < SPB; ++k) {
    const int s = s0 + min(k, ns - 1);
    sc[k] = k < ns ? a.src_cell[s] : -1;
    wv[k] = k < ns ? a.wav[(long long)s * a.wav_stride + a.it] : 0.f;
    so4[k] = (long long)s * p4;
  }
  const long long row4 = (long long)y * d.GX + g;  // (y, g) within a plane
  float4 vzq[SPB][4], pq[SPB][4];
#pragma unroll
  for (int k = 0; k < SPB; ++k) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int zz = zs - 2 + j;
      vzq[k][j] = (tin && k < ns && zz >= 0) ? A4[3 * f4 + so4[k] + zz * sz4 + row4] : z4();
      pq[k][j] = z4();
    }
    vzq[k][3] = z4();
    pq[k][3] = z4();
  }
  const long long inner = (long long)d.nz * d.ny * d.nx;
  float* eg = a.energy ? a.energy + (long long)grp * 4 * inner : nullptr;
  int par = 0;
  for (int zp = zs; zp <= zend; ++zp) {
    const bool hp = zp < d.NZ;
    float4 vxc[SPB], vyc[SPB];
    // phase A: p' of plane zp on the tile and its halo, into LDS
    if (hp && tin) {
      const long long q4 = zp * sz4 + row4;
      const float4 omk1 = C[q4], a1 = C[p4 + q4];
#pragma unroll
      for (int k = 0; k < SPB; ++k) {
        if (k >= ns) continue;
        const float4* P = A4 + so4[k];
        vzq[k][3] = zp + 1 < d.NZ ? P[3 * f4 + q4 + sz4] : z4();
        float4 ds = z4();
        const float4 pn = march_p<REC>(d, P, P + f4, P + 2 * f4, vzq[k], zp, y, g, omk1, a1,
                                       sc[k], wv[k], a.dt, vxc[k], vyc[k], ds);
        lp[k][par][ty + 1][gx + 1] = pn;
        pq[k][3] = pn;
        if (REC && zp >= zb && zp < ze) {
          reinterpret_cast<float4*>(a.dstore)[so4[k] + q4] = ds;
          if (a.pstore) reinterpret_cast<float4*>(a.pstore)[so4[k] + q4] = pn;
        }
      }
    }
    if (hp && hslot) {
      if (hin) {
        const long long q4 = (zp * (long long)d.NY + hy) * d.GX + hg;
        const float4 omk1 = C[q4], a1 = C[p4 + q4];
#pragma unroll
        for (int k = 0; k < SPB; ++k) {
          if (k >= ns) continue;
          const float4* P = A4 + so4[k];
          const float4* VZ = P + 3 * f4;
          float4 hv[4];
          hv[0] = zp >= 2 ? VZ[q4 - 2 * sz4] : z4();
          hv[1] = zp >= 1 ? VZ[q4 - sz4] : z4();
          hv[2] = VZ[q4];
          hv[3] = zp + 1 < d.NZ ? VZ[q4 + sz4] : z4();
          float4 hx, hyv, ds;
          lp[k][par][hr][hc] = march_p<false>(d, P, P + f4, P + 2 * f4, hv, zp, hy, hg, omk1, a1,
                                              sc[k], wv[k], a.dt, hx, hyv, ds);
        }
      } else {
#pragma unroll
        for (int k = 0; k < SPB; ++k) lp[k][par][hr][hc] = z4();
      }
    }
    if (!hp) {
#pragma unroll
      for (int k = 0; k < SPB; ++k) { vzq[k][3] = z4(); pq[k][3] = z4(); }
    }
    __syncthreads();
    // phase B: vx', vy' at zp; vz' at zp - 2
    if (tin) {
      const int zv = zp - 2;
      if (hp && zp >= zb && zp < ze) {
        const long long q4 = zp * sz4 + row4;
        const float4 omk2 = C[2 * p4 + q4], omk3 = C[3 * p4 + q4], a2 = C[5 * p4 + q4];
        unsigned mx = 0, my = 0;
        const bool zx = zp >= d.f && zp <= d.NZ - 2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int x = 4 * g + e;
          if (zx && y >= 1 && y <= d.NY - 2 && x >= 1 && x <= d.NX - 3) mx |= 1u << e;
          if (zx && y >= 1 && y <= d.NY - 3 && x >= 1 && x <= d.NX - 2) my |= 1u << e;
        }
        float4 ep = z4(), ex = z4(), ey = z4();
#pragma unroll
        for (int k = 0; k < SPB; ++k) {
          if (k >= ns) continue;
          const float4 pa = lp[k][par][ty + 1][gx], pb = lp[k][par][ty + 1][gx + 1];
          const float4 pc = lp[k][par][ty + 1][gx + 2];
          const float4 pu = lp[k][par][ty][gx + 1], pd1 = lp[k][par][ty + 2][gx + 1];
          const float4 pd2 = lp[k][par][ty + 3][gx + 1];
          const float ps[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w,
                                pc.x, pc.y, pc.z, pc.w};
          float4 vx = vxc[k], vy = vyc[k];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if ((mx >> e) & 1u) {
              const float u = g3(omk2, e) * g3(vx, e);
              s3(vx, e, u - g3(a2, e) * d4(ps[5 + e], ps[4 + e], ps[6 + e], ps[3 + e]));
            }
            if ((my >> e) & 1u) {
              const float u = g3(omk3, e) * g3(vy, e);
              s3(vy, e, u - g3(a2, e) * d4(g3(pd1, e), g3(pb, e), g3(pd2, e), g3(pu, e)));
            }
          }
          float4* Bk = B4 + so4[k];
          Bk[q4] = pb;
          Bk[f4 + q4] = vx;
          Bk[2 * f4 + q4] = vy;
          if (eg) {
            ep = make_float4(ep.x + pb.x * pb.x, ep.y + pb.y * pb.y, ep.z + pb.z * pb.z, ep.w + pb.w * pb.w);
            ex = make_float4(ex.x + vx.x * vx.x, ex.y + vx.y * vx.y, ex.z + vx.z * vx.z, ex.w + vx.w * vx.w);
            ey = make_float4(ey.x + vy.x * vy.x, ey.y + vy.y * vy.y, ey.z + vy.z * vy.z, ey.w + vy.w * vy.w);
          }
        }
        if (eg) {
          energy_add(d, eg, zp, y, g, ep);
          energy_add(d, eg + inner, zp, y, g, ex);
          energy_add(d, eg + 2 * inner, zp, y, g, ey);
        }
      }
      if (hp && fsm && zp == d.f + 1) {  // p[f-1] = -p'[f+1] (plane f+1 may be the next chunk's)
        const long long q4 = (long long)(d.f - 1) * sz4 + row4;
#pragma unroll
        for (int k = 0; k < SPB; ++k)
          if (k < ns)
            B4[so4[k] + q4] = make_float4(-pq[k][3].x, -pq[k][3].y, -pq[k][3].z, -pq[k][3].w);
      }
      const bool vin = zv >= zb && zv < ze;
      if (vin || (fsm && zv == d.f)) {
        const long long q4 = zv * sz4 + row4;
        const float4 omk4 = C[4 * p4 + q4], a2 = C[5 * p4 + q4];
        unsigned mz = 0;
        if (zv >= d.f && zv <= d.NZ - 3 && y >= 1 && y <= d.NY - 2)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int x = 4 * g + e;
            if (x >= 1 && x <= d.NX - 2) mz |= 1u << e;
          }
        const bool mirror = d.fs && zv == d.f;  // p[f-1] is read as -p[f+1]
        float4 ez = z4();
#pragma unroll
        for (int k = 0; k < SPB; ++k) {
          if (k >= ns) continue;
          float4 vz = vzq[k][0];
          if (mz) {
            const float4 pm = mirror ? make_float4(-pq[k][2].x, -pq[k][2].y, -pq[k][2].z, -pq[k][2].w)
                                     : pq[k][0];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              if (!((mz >> e) & 1u)) continue;
              const float u = g3(omk4, e) * g3(vz, e);
              s3(vz, e, u - g3(a2, e) * d4(g3(pq[k][2], e), g3(pq[k][1], e), g3(pq[k][3], e),
                                            g3(pm, e)));
            }
          }
          float4* Bk = B4 + so4[k];
          if (vin) Bk[3 * f4 + q4] = vz;
          if (fsm && zv == d.f) Bk[3 * f4 + q4 - sz4] = vz;
          if (eg && vin) ez = make_float4(ez.x + vz.x * vz.x, ez.y + vz.y * vz.y, ez.z + vz.z * vz.z, ez.w + vz.w * vz.w);
        }
        if (eg && vin) energy_add(d, eg + 3 * inner, zv, y, g, ez);
      }
    }
#pragma unroll
    for (int k = 0; k < SPB; ++k) {
      vzq[k][0] = vzq[k][1]; vzq[k][1] = vzq[k][2]; vzq[k][2] = vzq[k][3];
      pq[k][0] = pq[k][1]; pq[k][1] = pq[k][2]; pq[k][2] = pq[k][3];
    }
    par ^= 1;
  }
}

// ---------------------------------------------------------------------------
// Fused step, rows in wavefronts (launch choice step_kernel = 2).  A workgroup
// of 16 wavefronts owns a 61-column x 13-row (x, y) window of one z segment
// for every shot of a group (at most 8).  Lane l of wavefront w holds the cell
// (y0 - 1 + w, xs + l) of each shot, and the workgroup marches down z.  Per
// plane z:
//   P: p' of plane z from the old state: vx of the neighbouring columns from
//      the adjacent lanes (DPP wavefront shifts; the window's outer columns
//      load theirs), vy of the rows above and below from memory (the rows of
//      the neighbouring wavefronts: L1 / L2), vz of planes z-2 .. z+1 from a
//      register queue; the source term; then vz' of plane z - 2, whose stencil
//      is p' of planes z-3 .. z (a register queue of p'); p' of plane z to LDS;
//   barrier;
//   V: vx', vy' of plane z: p' of the neighbouring columns by DPP, of the rows
//      above and below from LDS.
// Each field is read from HBM once per step and written once; the energy sums
// run over the group's shots in shot order with one read-modify-write of the
// group's partials per cell.  The state is out of place (A -> B): the window's
// halo cells are recomputed from A while neighbouring windows write B.
// Owned cells: lanes 1..61 (vx' reads p' two columns right), wavefronts 1..13
// (vy' reads p' two rows down); every other lane and wavefront is halo.
// Arithmetic per cell is p3_step_kernel's and v3_step_kernel's, in the same
// order, so the step is bit-identical to them.
// ---------------------------------------------------------------------------
constexpr int F3_NW = 16;     // wavefronts (rows) per workgroup
constexpr int F3_OWN_X = 61;  // owned columns per window (lanes 1..61)
constexpr int F3_OWN_Y = 13;  // owned rows per window (wavefronts 1..13)

struct Fused3Args {
  Dims3 d;
  int S, it, nt, spb;
  float dt;
  const float* A;           // state before the step, [field][shot][plane]
  float* B;                 // state after the step
  long long fstride;        // S * plane
  const float* coef;
  const int* src_cell;
  const float* wav; long long wav_stride;
  float* energy;            // [group][4][nz][ny][nx] or null
  int strips, ytiles, groups, seg_h;
  long long units;
};

// per-lane column flags of the fused step (bits of one word)
enum : unsigned {
  F3_OWN = 1u, F3_CP = 2u, F3_CVX = 4u, F3_CVY = 8u, F3_CE = 16u, F3_XIN = 32u,
  F3_L0 = 64u, F3_L01 = 128u, F3_L63 = 256u
};

// the loads of one shot at one plane (issued ahead of their use)
struct F3Load {
  float p, vx, vxh, vy, vym2, vym1, vyp1, vzn;
};

// Buffer accesses: a 128-bit resource (base, size in bytes) in scalar
// registers, a 32-bit lane byte offset and a uniform byte offset (the shot's
// plane).  Reads beyond the size return 0 and writes beyond it are dropped.
typedef __amdgpu_buffer_rsrc_t f3_rsrc;
__device__ __forceinline__ f3_rsrc f3_buf(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
// (the b32 builtins move raw 32-bit words: bit casts, not conversions)
__device__ __forceinline__ float f3_bld(f3_rsrc r, unsigned voff, unsigned soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void f3_bst(f3_rsrc r, unsigned voff, unsigned soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

template <int SPB>
__global__ __launch_bounds__(F3_NW * 64) void fwd3f_kernel(Fused3Args a) {
  __shared__ float lp[2][SPB][F3_NW][64];  // p' of the plane: parity, shot, row, lane
  // vz of planes z-2 .. z+1 of each thread's own cell (a ring by plane; each
  // thread reads only what it wrote, so the ring needs no barrier)
  __shared__ float vzr[3][SPB][F3_NW][64];
  const Dims3 d = a.d;
  const long long L = (long long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (L >= a.units) return;  // uniform per workgroup
  const int grp = (int)(L % a.groups);
  long long t = L / a.groups;
  const int strip = (int)(t % a.strips);
  t /= a.strips;
  const int ytile = (int)(t % a.ytiles);
  const int zseg = (int)(t / a.ytiles);
  // segment bounds; no boundary at f, so planes f-1 and f (the mirrors) share a segment
  auto zbound = [&](int i) {
    int b = min(d.NZ, i * a.seg_h);
    if (d.fs && b == d.f) b = min(d.NZ, d.f + 1);
    return b;
  };
  const int za = zbound(zseg), zb = zbound(zseg + 1);
  if (za >= zb) return;
  const int s0 = grp * a.spb, ns = min(min(SPB, a.spb), a.S - s0);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x = strip * F3_OWN_X - 1 + lane;
  const int y = ytile * F3_OWN_Y - 1 + w;  // uniform per wavefront
  const unsigned xc = (unsigned)min(max(x, 0), d.PX - 1);
  // the vx the window's outer lanes need from outside it: lanes 0, 1 load
  // column x - 2, lane 63 column x + 1
  const unsigned xh = (unsigned)min(max(lane <= 1 ? x - 2 : (lane == 63 ? x + 1 : x), 0), d.PX - 1);
  const bool xin = x >= 0 && x < d.NX, yin = y >= 0 && y < d.NY;
  unsigned cflags = 0;
  cflags |= (lane >= 1 && lane <= F3_OWN_X && xin) ? F3_OWN : 0u;
  cflags |= (x >= 2 && x <= d.NX - 3) ? F3_CP : 0u;
  cflags |= (x >= 1 && x <= d.NX - 3) ? F3_CVX : 0u;
  cflags |= (x >= 1 && x <= d.NX - 2) ? F3_CVY : 0u;  // vy and vz
  cflags |= (x >= d.nabc && x < d.nabc + d.nx) ? F3_CE : 0u;
  cflags |= xin ? F3_XIN : 0u;
  cflags |= lane == 0 ? F3_L0 : 0u;
  cflags |= lane <= 1 ? F3_L01 : 0u;
  cflags |= lane == 63 ? F3_L63 : 0u;
  const bool own_row = w >= 1 && w <= F3_OWN_Y && yin;  // uniform
  const bool rp = y >= 2 && y <= d.NY - 3;
  const bool rvx = y >= 1 && y <= d.NY - 2, rvy = y >= 1 && y <= d.NY - 3;  // rvx: vx and vz
  const bool re = y >= d.nabc && y < d.nabc + d.ny;
  auto yclamp = [&](int v) { return min(max(v, 0), d.NY - 1); };
  const unsigned ry = (unsigned)yclamp(y), rym2 = (unsigned)yclamp(y - 2),
                 rym1 = (unsigned)yclamp(y - 1), ryp1 = (unsigned)yclamp(y + 1);
  const unsigned PX = (unsigned)d.PX, NY = (unsigned)d.NY;
  auto zc = [&](int z) { return (unsigned)min(max(z, 0), d.NZ - 1); };
  // byte offset of (z, row, column) in a plane
  auto bo = [&](unsigned z, unsigned row, unsigned col) { return ((z * NY + row) * PX + col) * 4u; };
  const unsigned PB = (unsigned)(d.plane * 4);  // bytes of one plane (the host checks the sizes)
  const long long gbase = (long long)s0 * d.plane;
  const unsigned gbytes = (unsigned)ns * PB;
  const f3_rsrc rP = f3_buf(a.A + gbase, gbytes), rVX = f3_buf(a.A + a.fstride + gbase, gbytes),
                rVY = f3_buf(a.A + 2 * a.fstride + gbase, gbytes),
                rVZ = f3_buf(a.A + 3 * a.fstride + gbase, gbytes);
  const f3_rsrc wP = f3_buf(a.B + gbase, gbytes), wVX = f3_buf(a.B + a.fstride + gbase, gbytes),
                wVY = f3_buf(a.B + 2 * a.fstride + gbase, gbytes),
                wVZ = f3_buf(a.B + 3 * a.fstride + gbase, gbytes);
  const f3_rsrc rC = f3_buf(a.coef, 6u * PB);
  const long long inner = (long long)d.nz * d.ny * d.nx;
  const unsigned IB = (unsigned)(inner * 4);
  const bool has_e = a.energy != nullptr;
  const f3_rsrc rE = f3_buf(has_e ? a.energy + (long long)grp * 4 * inner : a.coef,
                            has_e ? 4u * IB : 0u);
  const bool erow = has_e && own_row && re;  // uniform
  const unsigned eb = (unsigned)(((y - d.nabc) * d.nx + (x - d.nabc)) * 4);  // owned interior only
  const unsigned EPB = (unsigned)(d.ny * d.nx * 4);  // bytes of one energy plane
  const int cell0 = y * d.PX + x;  // cell index within a plane (valid rows and columns)
  // per shot: source cell, source term of this step
  int sc[SPB];
  float wv[SPB];
#pragma unroll
  for (int k = 0; k < SPB; ++k) {
    const int s = s0 + min(k, ns - 1);
    sc[k] = k < ns ? a.src_cell[s] : -1;
    wv[k] = k < ns ? a.wav[(long long)s * a.wav_stride + a.it] : 0.f;
  }
  // shot k's plane offset (inactive shots read the group's last shot)
  auto soff = [&](int k) { return (unsigned)min(k, ns - 1) * PB; };
  // the loads of shot k at plane z
  auto issue = [&](int k, int z) {
    const unsigned zz = zc(z), so = soff(k);
    const unsigned o = bo(zz, ry, xc);
    F3Load v;
    v.p = f3_bld(rP, o, so);
    v.vx = f3_bld(rVX, o, so);
    v.vxh = f3_bld(rVX, bo(zz, ry, xh), so);
    v.vy = f3_bld(rVY, o, so);
    v.vym2 = f3_bld(rVY, bo(zz, rym2, xc), so);
    v.vym1 = f3_bld(rVY, bo(zz, rym1, xc), so);
    v.vyp1 = f3_bld(rVY, bo(zz, ryp1, xc), so);
    v.vzn = f3_bld(rVZ, bo(zc(z + 1), ry, xc), so);
    return v;
  };
  // queues: vz of planes z-2, z-1, z (old state) in the LDS ring, slot
  // (plane + 3) % 3; p' of planes z-3, z-2, z-1 in registers
  float pq0[SPB], pq1[SPB], pq2[SPB], hvx[SPB], hvy[SPB];
  auto slot = [](int z) { return (z + 3) % 3; };
#pragma unroll
  for (int k = 0; k < SPB; ++k) {
    vzr[slot(za - 3)][k][w][lane] = f3_bld(rVZ, bo(zc(za - 3), ry, xc), soff(k));
    vzr[slot(za - 2)][k][w][lane] = f3_bld(rVZ, bo(zc(za - 2), ry, xc), soff(k));
    vzr[slot(za - 1)][k][w][lane] = f3_bld(rVZ, bo(zc(za - 1), ry, xc), soff(k));
    pq0[k] = pq1[k] = pq2[k] = 0.f;
    hvx[k] = hvy[k] = 0.f;
  }
  float a2q0 = f3_bld(rC, bo(zc(za - 3), ry, xc), 5u * PB);  // a2 of plane z-2
  float a2q1 = f3_bld(rC, bo(zc(za - 2), ry, xc), 5u * PB);  // a2 of plane z-1
  // loads in flight: shots 0 and 1 of the first plane
  F3Load q0 = issue(0, za - 1);
  F3Load q1 = issue(SPB > 1 ? 1 : 0, SPB > 1 ? za - 1 : za);
  int par = 0;
  for (int z = za - 1; z <= zb + 1; ++z) {
    unsigned cf = cflags;
    asm volatile("" : "+v"(cf));  // re-derive the lane tests each plane (no SGPR masks kept)
    const bool zown = z >= za && z < zb;  // uniform: this plane's p, vx, vy are stored
    const int zv = z - 2;                 // the plane whose vz' this iteration completes
    const bool vown = zv >= za && zv < zb;
    const unsigned o = bo(zc(z), ry, xc);
    const unsigned ov = bo(zc(zv), ry, xc);
    const bool ez_row = erow && vown && zv >= d.nabc && zv < d.nabc + d.nz;
    const bool e_on = erow && zown && z >= d.nabc && z < d.nabc + d.nz;  // uniform
    const bool ecell = (cf & F3_OWN) && (cf & F3_CE);
    // every load of the plane that is not a shot's state, issued before the
    // shots' loads so that waiting for it never waits for those: the
    // coefficients and the energy partials (read at a valid offset when unused)
    const float omk1 = f3_bld(rC, o, 0u), a1 = f3_bld(rC, o, PB);
    const float omk4v = f3_bld(rC, ov, 4u * PB);
    const float a2z = f3_bld(rC, o, 5u * PB);
    const float omk2 = f3_bld(rC, o, 2u * PB), omk3 = f3_bld(rC, o, 3u * PB);
    const unsigned oez = (ez_row && ecell) ? (unsigned)(zv - d.nabc) * EPB + eb : 0u;
    const unsigned oe = (e_on && ecell) ? (unsigned)(z - d.nabc) * EPB + eb : 0u;
    const float ez0 = f3_bld(rE, oez, 3u * IB);
    const float ep0 = f3_bld(rE, oe, 0u), ex0 = f3_bld(rE, oe, IB), ey0 = f3_bld(rE, oe, 2u * IB);
    const bool zP = z >= d.f + 1 && z <= d.NZ - 3 && rp;  // uniform parts of the regions
    const bool zVZ = zv >= d.f && zv <= d.NZ - 3 && rvx;
    const bool zvalid = yin && z >= 0 && z < d.NZ;
    const int cell = z * d.NY * d.PX + cell0;
    const bool mirror_vz = d.fs && zv == d.f;  // vz' of plane f reads p'(f-1) = -p'(f+1)
    const bool st_row = own_row;
    float ez = 0.f;
    const int s_m2 = slot(z - 2), s_m1 = slot(z - 1), s_00 = slot(z);
    // ---- phase P ----
#pragma unroll
    for (int k = 0; k < SPB; ++k) {
      const float vzm2 = vzr[s_m2][k][w][lane], vzm1 = vzr[s_m1][k][w][lane],
                  vz00 = vzr[s_00][k][w][lane];
      const F3Load v = q0;
      q0 = q1;
      // two (plane, shot) pairs ahead
      q1 = issue((k + 2) % SPB, z + (k + 2) / SPB);
      const bool act = k < ns;  // uniform
      const float l1 = lane_prev(v.vx), l2 = lane_prev(l1), r1 = lane_next(v.vx), hn = lane_next(v.vxh);
      const float vm1 = (cf & F3_L0) ? hn : l1;
      const float vm2 = (cf & F3_L01) ? v.vxh : l2;
      const float vp1 = (cf & F3_L63) ? v.vxh : r1;
      const float dvx = d4(v.vx, vm1, vp1, vm2);
      const float dvy = d4(v.vy, v.vym1, v.vyp1, v.vym2);
      const float dvz = d4(vz00, vzm1, v.vzn, vzm2);
      float sum = dvx + dvy;
      sum = sum + dvz;
      const float u = omk1 * v.p;
      const float pt = u - a1 * sum;
      float pn = (zP && (cf & F3_CP)) ? pt : v.p;
      if (zvalid && (cf & F3_XIN) && cell == sc[k]) pn = pn + a.dt * wv[k];
      // vz' of plane z-2: p' of planes z-3 (pq0), z-2 (pq1), z-1 (pq2), z (pn)
      const float pu = mirror_vz ? -pq2[k] : pq0[k];
      const float vzo = vzm2;
      const float uz = omk4v * vzo;
      const float vzt = uz - a2q0 * d4(pq2[k], pq1[k], pn, pu);
      const float vzf = (zVZ && (cf & F3_CVY)) ? vzt : vzo;
      if (act) {
        if (zown) lp[par][k][w][lane] = pn;
        const unsigned so = (unsigned)k * PB;
        if (st_row && (cf & F3_OWN)) {
          if (vown) {
            if (!(d.fs && zv == d.f - 1)) f3_bst(wVZ, ov, so, vzf);
            if (mirror_vz) f3_bst(wVZ, ov - NY * PX * 4u, so, vzf);  // vz[f-1] = vz[f]
            if (ez_row) ez = ez + vzf * vzf;
          }
          if (zown && !(d.fs && z == d.f - 1)) f3_bst(wP, o, so, pn);
          if (d.fs && z == d.f + 1 && d.f - 1 >= za && d.f - 1 < zb)
            f3_bst(wP, o - 2u * NY * PX * 4u, so, -pn);  // p[f-1] = -p[f+1]
        }
      }
      vzr[s_m2][k][w][lane] = v.vzn;  // plane z+1 takes plane z-2's slot
      pq0[k] = pq1[k]; pq1[k] = pq2[k]; pq2[k] = pn;
      hvx[k] = v.vx; hvy[k] = v.vy;
    }
    if (ez_row && ecell) f3_bst(rE, oez, 3u * IB, ez0 + ez);
    a2q0 = a2q1;
    a2q1 = a2z;
    if (zown) {
      __syncthreads();
      // ---- phase V ----
      if (own_row) {
        const bool zv_in = z >= d.f && z <= d.NZ - 2;
        const bool inVX = zv_in && rvx && (cf & F3_CVX), inVY = zv_in && rvy && (cf & F3_CVY);
        float ep = 0.f, ex = 0.f, ey = 0.f;
#pragma unroll
        for (int k = 0; k < SPB; ++k) {
          if (k >= ns) break;  // uniform
          const unsigned so = (unsigned)k * PB;
          const float pc = pq2[k];
          const float pm1 = lane_prev(pc), pp1 = lane_next(pc), pp2 = lane_next(pp1);
          const float pym1 = lp[par][k][w - 1][lane], pyp1 = lp[par][k][w + 1][lane],
                      pyp2 = lp[par][k][w + 2][lane];
          const float ux = omk2 * hvx[k];
          const float tvx = ux - a2z * d4(pp1, pc, pp2, pm1);
          const float uy = omk3 * hvy[k];
          const float tvy = uy - a2z * d4(pyp1, pc, pyp2, pym1);
          const float vxf = inVX ? tvx : hvx[k];
          const float vyf = inVY ? tvy : hvy[k];
          if (cf & F3_OWN) {
            f3_bst(wVX, o, so, vxf);
            f3_bst(wVY, o, so, vyf);
          }
          if (e_on) {
            ep = ep + pc * pc;
            ex = ex + vxf * vxf;
            ey = ey + vyf * vyf;
          }
        }
        if (e_on && ecell) {
          f3_bst(rE, oe, 0u, ep0 + ep);
          f3_bst(rE, oe, IB, ex0 + ex);
          f3_bst(rE, oe, 2u * IB, ey0 + ey);
        }
      }
      par ^= 1;
    }
  }
}

// shared wavelet: grad[it] = sum over shots of the shot's source term, in shot order
// several points per shot, one wavelet per shot: the K points' samples of
// each (shot, step) summed in point order
__global__ void wavelet_points3_kernel(const float* __restrict__ gwp, int S, int K, int nt,
                                       float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)S * nt) return;
  const long long s = t / nt, i = t - s * nt;
  float v = gwp[s * K * nt + i];
  for (int k = 1; k < K; ++k) v = v + gwp[(s * K + k) * nt + i];
  out[t] = v;
}

__global__ void wavelet_sum3_kernel(const float* gws, int S, float* gw, int it) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float v = 0.f;
  for (int s = 0; s < S; ++s) v = v + gws[s];
  gw[it] = v;
}

// chain rule through alpha1 = den c c dt / dz and alpha2 = dt / (den dz), and
// the transpose of pad_model_3d: each model cell gathers the padded cells that
// copy it (edges and corners collect whole blocks of the absorbing layers)
__global__ void fold3_kernel(Dims3 d, int groups, const float* gA1, const float* gA2,
                             const float* vp, const float* rho, float dt, float dz,
                             float* gvp, float* grho) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (long long)d.nz * d.ny * d.nx) return;
  const int e = (int)(m % d.nx);
  const int b = (int)((m / d.nx) % d.ny);
  const int a = (int)(m / ((long long)d.nx * d.ny));
  const int i0 = a == 0 ? 0 : a + d.nabc, i1 = a == d.nz - 1 ? d.NZ - 1 : a + d.nabc;
  const int j0 = b == 0 ? 0 : b + d.nabc, j1 = b == d.ny - 1 ? d.NY - 1 : b + d.nabc;
  const int k0 = e == 0 ? 0 : e + d.nabc, k1 = e == d.nx - 1 ? d.NX - 1 : e + d.nabc;
  const float vc = vp[m], den = rho[m];
  float f1 = 2.f * den;
  f1 = f1 * vc;
  f1 = f1 * dt;
  f1 = f1 / dz;
  float r1 = vc * vc;
  r1 = r1 * dt;
  r1 = r1 / dz;
  float r2 = den * den;
  r2 = r2 * dz;
  r2 = dt / r2;
  float av = 0.f, ar = 0.f;
  for (int i = i0; i <= i1; ++i)
    for (int j = j0; j <= j1; ++j)
      for (int k = k0; k <= k1; ++k) {
        const long long q = ((long long)i * d.NY + j) * d.PX + k;
        float A1 = gA1[q], A2 = gA2 ? gA2[q] : 0.f;
        for (int gi = 1; gi < groups; ++gi) {
          A1 = A1 + gA1[(long long)gi * d.plane + q];
          if (gA2) A2 = A2 + gA2[(long long)gi * d.plane + q];
        }
        av = av + A1 * f1;
        if (grho) ar = ar + (A1 * r1 - A2 * r2);
      }
  if (gvp) gvp[m] = av;
  if (grho) grho[m] = ar;
}

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

struct Plan3 {
  Dims3 d;
  int S, R, nt, spb, groups, echunk, k, nseg;
  int K;                            // source points per shot
  bool march;                       // fused z-marching step (state in and out of place)
  bool fused;                       // fused step with rows in wavefronts (state out of place)
  int tiles_x, tiles_y, zc;         // z-marching decomposition
  long long units;
  int f_strips, f_ytiles;           // fused step: 61 x 13 windows
  size_t off_coef, off_state, off_E, off_src, off_rcell;
  size_t off_lam, off_r, off_D, off_P1, off_gA1, off_gA2, off_gws, off_ucell, off_ustart,
      off_urcv, off_rowstart, off_gwp;
  size_t total;
};

// pass 0: forward; 1: backward; 2: backward with the density gradient
int make_plan3(const gbw_acoustic3d_problem* p, int pass, Plan3& pl) {
  if (!p) return fail(GBW_ERR_INVALID, "problem is NULL");
  if (pass < 0 || pass > 2) return fail(GBW_ERR_INVALID, "pass must be 0, 1 or 2");
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
    return fail(GBW_ERR_INVALID, "Grid dimensions must be positive");
  if (p->nabc < 1) return fail(GBW_ERR_INVALID, "nabc must be >= 1");
  if (p->nt <= 0) return fail(GBW_ERR_INVALID, "nt must be positive");
  if (p->src_n <= 0) return fail(GBW_ERR_INVALID, "src_n must be positive");
  if (p->rcv_n < 0) return fail(GBW_ERR_INVALID, "rcv_n must be >= 0");
  if (!(p->dt > 0.f) || !(p->dz > 0.f)) return fail(GBW_ERR_INVALID, "dt and dz must be positive");
  if (p->checkpoint_every < 0) return fail(GBW_ERR_INVALID, "checkpoint_every must be >= 0");
  if (pass > 0 && p->checkpoint_every == 0)
    return fail(GBW_ERR_INVALID, "the gradient needs checkpoint_every > 0");
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
  // in-volume float4 accesses use 32-bit byte offsets
  if (d.plane >= (1LL << 30)) return fail(GBW_ERR_INVALID, "padded volume exceeds 2^30 cells");
  pl.S = p->src_n; pl.R = p->rcv_n; pl.nt = p->nt;
  if (p->step_kernel < 0 || p->step_kernel > 2)
    return fail(GBW_ERR_INVALID, "step_kernel must be 0, 1 or 2");
  if (p->src_points < 0) return fail(GBW_ERR_INVALID, "src_points must be >= 0");
  pl.K = std::max(1, p->src_points);
  if (p->wavelet_shared < 0 || p->wavelet_shared > 2)
    return fail(GBW_ERR_INVALID, "wavelet_shared must be 0, 1 or 2");
  if (pl.K > 1 && p->wavelet_shared == 1)
    return fail(GBW_ERR_INVALID,
                "several source points per shot take src_v of shape (src_n, nt) or "
                "(src_n, src_points, nt)");
  if (p->z_chunk < 0) return fail(GBW_ERR_INVALID, "z_chunk must be >= 0");
  // several points per shot: the p / v kernels (the fused steps inject one)
  pl.march = p->step_kernel == 1 && pl.K == 1;
  // the fused step runs the forward pass; the gradient's recompute keeps the
  // in-place p / v kernels
  pl.fused = p->step_kernel == 2 && pass == 0 && pl.K == 1;
  int spb;
  if (pl.fused) {
    // its buffer accesses take 32-bit byte offsets: the six coefficient planes
    // and a group's planes of one field must each stay below 4 GiB
    const long long pb = d.plane * 4;
    if (6 * pb > 0xFFFFFFFFLL)
      return fail(GBW_ERR_INVALID, "the fused step needs a padded volume below 2^28.4 cells");
    const int most = (int)std::min<long long>(8, 0xFFFFFFFFLL / pb);
    spb = p->shots_per_block > 0 ? std::min(p->shots_per_block, pl.S) : std::min(pl.S, most);
    if (spb > most)
      return fail(GBW_ERR_INVALID, "shots_per_block must be at most " + std::to_string(most) +
                                       " with the fused step on this grid");
  } else if (pl.march) {
    spb = p->shots_per_block > 0 ? p->shots_per_block : 4;
    if (spb != 1 && spb != 2 && spb != 4)
      return fail(GBW_ERR_INVALID, "shots_per_block must be 0, 1, 2 or 4 with the z-marching kernel");
  } else {
    spb = p->shots_per_block > 0 ? std::min(p->shots_per_block, pl.S) : pl.S;
  }
  pl.spb = spb;
  pl.groups = (pl.S + spb - 1) / spb;
  if (pl.groups > 65535) return fail(GBW_ERR_INVALID, "more than 65535 shot groups: raise shots_per_block");
  pl.tiles_x = d.GX / 16;
  pl.tiles_y = (d.NY + 15) / 16;
  {
    // about eight workgroups per CU, each marching at least 8 planes
    const long long per_z = (long long)pl.tiles_x * pl.tiles_y * pl.groups;
    long long nzc = std::max<long long>(1, (2048 + per_z / 2) / per_z);
    nzc = std::min<long long>(nzc, std::max(1, d.NZ / 8));
    pl.zc = p->z_chunk > 0 ? std::min(p->z_chunk, d.NZ) : (int)((d.NZ + nzc - 1) / nzc);
    pl.units = per_z * ((d.NZ + pl.zc - 1) / pl.zc);
  }
  pl.f_strips = (d.NX + F3_OWN_X - 1) / F3_OWN_X;
  pl.f_ytiles = (d.NY + F3_OWN_Y - 1) / F3_OWN_Y;
  pl.echunk = p->energy_chunk > 0 ? std::min(p->energy_chunk, p->nt) : p->nt;
  pl.k = p->checkpoint_every > 0 ? std::min(p->checkpoint_every, p->nt) : 0;
  pl.nseg = pl.k > 0 ? (pl.nt + pl.k - 1) / pl.k : 0;
  const size_t inner = (size_t)d.nz * d.ny * d.nx;
  const size_t fld = (size_t)pl.S * d.plane * sizeof(float);  // one field, all shots
  const size_t R1 = (size_t)std::max(1, pl.R);
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += al256(bytes); return at; };
  pl.off_coef = take(6 * (size_t)d.plane * sizeof(float));
  pl.off_state = take((pl.march || pl.fused ? 8 : 4) * fld);
  pl.off_src = take((size_t)pl.S * pl.K * sizeof(int));
  pl.off_rcell = take(R1 * sizeof(int));
  pl.off_gwp = 0;
  if (pass == 0) {
    pl.off_E = take((size_t)pl.groups * 4 * inner * sizeof(float));
  } else {
    pl.off_lam = take(4 * fld);
    pl.off_r = take(fld);
    pl.off_D = take((size_t)pl.k * fld);
    pl.off_P1 = pass == 2 ? take((size_t)pl.k * fld) : 0;
    pl.off_gA1 = take((size_t)pl.groups * d.plane * sizeof(float));
    pl.off_gA2 = pass == 2 ? take((size_t)pl.groups * d.plane * sizeof(float)) : 0;
    pl.off_gws = take((size_t)pl.S * sizeof(float));
    pl.off_ucell = take(R1 * sizeof(int));
    pl.off_ustart = take((R1 + 1) * sizeof(int));
    pl.off_urcv = take(R1 * sizeof(int));
    pl.off_rowstart = take(((size_t)d.NZ * d.NY + 1) * sizeof(int));
    // several points per shot: wavelet-gradient samples per point
    if (pl.K > 1) pl.off_gwp = take((size_t)pl.S * pl.K * pl.nt * sizeof(float));
  }
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

// source and receiver cells, with the reference's IndexError semantics (vx
// has NX-1 columns, vy NY-1 rows, vz NZ-1 planes, and every receiver reads all four)
int geometry3(const Plan3& pl, const gbw_acoustic3d_inputs* in, std::vector<int>& src,
              std::vector<int>& rcell) {
  const Dims3& d = pl.d;
  if (!in->vp || !in->rho || !in->damp || !in->src_v)
    return fail(GBW_ERR_INVALID, "vp, rho, damp and src_v must be device pointers");
  if (!in->src_x || !in->src_y || !in->src_z) return fail(GBW_ERR_INVALID, "source positions are NULL");
  if (pl.R > 0 && (!in->rcv_x || !in->rcv_y || !in->rcv_z))
    return fail(GBW_ERR_INVALID, "receiver positions are NULL");
  // point k of shot s at s * K + k
  src.assign((size_t)pl.S * pl.K, 0);
  rcell.assign(std::max(1, pl.R), 0);
  for (int i = 0; i < pl.S * pl.K; ++i) {
    const long long z = in->src_z[i] + d.nabc, y = in->src_y[i] + d.nabc, x = in->src_x[i] + d.nabc;
    const std::string who = pl.K > 1 ? "source " + std::to_string(i / pl.K) + " point " +
                                           std::to_string(i % pl.K)
                                     : "source " + std::to_string(i);
    if (z < 0 || z >= d.NZ || y < 0 || y >= d.NY || x < 0 || x >= d.NX)
      return fail(GBW_ERR_INDEX, "index out of range: " + who);
    src[i] = (int)((z * d.NY + y) * d.PX + x);
    // the reference's indexed update keeps one of two equal cells' sums
    // (acoustic_3d.py:174): such shots are refused, not guessed at
    for (int j = i - i % pl.K; j < i; ++j)
      if (src[j] == src[i])
        return fail(GBW_ERR_INVALID, who + " repeats the cell of point " +
                                         std::to_string(j % pl.K) + " of its shot");
  }
  for (int r = 0; r < pl.R; ++r) {
    const long long z = in->rcv_z[r] + d.nabc, y = in->rcv_y[r] + d.nabc, x = in->rcv_x[r] + d.nabc;
    if (z < 0 || z >= d.NZ - 1 || y < 0 || y >= d.NY - 1 || x < 0 || x >= d.NX - 1)
      return fail(GBW_ERR_INDEX, "index out of range: receiver " + std::to_string(r));
    rcell[r] = (int)((z * d.NY + y) * d.PX + x);
  }
  return 0;
}

// workgroups of a step kernel: 8 x those of the largest y-slab
unsigned step_blocks3(const Dims3& d) {
  const long long slab = (long long)d.NZ * ((d.NY + 7) / 8) * d.GX;
  return (unsigned)(8 * ((slab + T3 - 1) / T3));
}

// one forward step of every shot (p, then v)
int forward_step3(Step3Args& a, const dim3& grid, hipStream_t st) {
  {
    LaunchTimer timer("p3_step_kernel", st);
    if (a.src_k > 1)
      p3_step_kernel<true><<<grid, T3, 0, st>>>(a);
    else
      p3_step_kernel<false><<<grid, T3, 0, st>>>(a);
  }
  LAUNCH_CHECK();
  {
    LaunchTimer timer("v3_step_kernel", st);
    v3_step_kernel<<<grid, T3, 0, st>>>(a);
  }
  LAUNCH_CHECK();
  return 0;
}

// one z-marching step: state A -> B
int march_step3(const Plan3& pl, March3Args& m, bool rec, hipStream_t st) {
  m.tiles_x = pl.tiles_x; m.tiles_y = pl.tiles_y; m.zc = pl.zc; m.groups = pl.groups;
  m.units = pl.units;
  const unsigned blocks = (unsigned)((pl.units + 7) / 8 * 8);
  LaunchTimer timer(rec ? "fwd3_march_kernel<recompute>" : "fwd3_march_kernel<forward>", st);
  switch (pl.spb * 2 + (rec ? 1 : 0)) {
    case 2: fwd3_march_kernel<1, false><<<blocks, MT3, 0, st>>>(m); break;
    case 3: fwd3_march_kernel<1, true><<<blocks, MT3, 0, st>>>(m); break;
    case 4: fwd3_march_kernel<2, false><<<blocks, MT3, 0, st>>>(m); break;
    case 5: fwd3_march_kernel<2, true><<<blocks, MT3, 0, st>>>(m); break;
    case 8: fwd3_march_kernel<4, false><<<blocks, MT3, 0, st>>>(m); break;
    case 9: fwd3_march_kernel<4, true><<<blocks, MT3, 0, st>>>(m); break;
    default: return fail(GBW_ERR_INVALID, "unsupported shots_per_block");
  }
  LAUNCH_CHECK();
  return 0;
}

// z segments of the fused step: the count that best fills the GPU's compute
// units in whole rounds of resident workgroups (one per CU), discounted by the
// three planes of p' each segment recomputes at its ends
int fused_segments(const Plan3& pl, int z_chunk, int cus) {
  const int NZ = pl.d.NZ;
  if (z_chunk > 0) return std::max(1, (NZ + z_chunk - 1) / z_chunk);
  const long long cols = (long long)pl.f_strips * pl.f_ytiles * pl.groups;
  int best = 1;
  double best_eff = 0.0;
  for (int zs = 1; zs <= std::max(1, NZ / 8); ++zs) {
    const int h = (NZ + zs - 1) / zs;
    const long long units = cols * ((NZ + h - 1) / h);
    const long long rounds = (units + cus - 1) / cus;
    const double eff = (double)units / (double)(rounds * cus) * h / (h + 3.0);
    if (eff > best_eff + 1e-9) { best_eff = eff; best = zs; }
  }
  return best;
}

// the fused step's decomposition (once per call: the current device's CU count)
void fused_setup3(const Plan3& pl, Fused3Args& f, int z_chunk) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      cus = v;
  }
  const int zs = fused_segments(pl, z_chunk, cus);
  f.seg_h = (pl.d.NZ + zs - 1) / zs;
  f.strips = pl.f_strips; f.ytiles = pl.f_ytiles; f.groups = pl.groups; f.spb = pl.spb;
  f.units = (long long)pl.f_strips * pl.f_ytiles * pl.groups * ((pl.d.NZ + f.seg_h - 1) / f.seg_h);
}

// one fused step: state A -> B
int fused_step3(const Plan3& pl, const Fused3Args& f, hipStream_t st) {
  const unsigned blocks = (unsigned)((f.units + 7) / 8 * 8);
  LaunchTimer timer("fwd3f_kernel", st);
  // the instance holds the next power of two of the group's shots
  if (pl.spb <= 1) fwd3f_kernel<1><<<blocks, F3_NW * 64, 0, st>>>(f);
  else if (pl.spb <= 2) fwd3f_kernel<2><<<blocks, F3_NW * 64, 0, st>>>(f);
  else if (pl.spb <= 4) fwd3f_kernel<4><<<blocks, F3_NW * 64, 0, st>>>(f);
  else fwd3f_kernel<8><<<blocks, F3_NW * 64, 0, st>>>(f);
  LAUNCH_CHECK();
  return 0;
}

}  // namespace

extern "C" {

size_t gbw_acoustic3d_workspace_bytes(const gbw_acoustic3d_problem* prob, int pass) {
  Plan3 pl;
  if (make_plan3(prob, pass, pl)) return 0;
  return pl.total;
}

size_t gbw_acoustic3d_checkpoint_bytes(const gbw_acoustic3d_problem* prob) {
  Plan3 pl;
  if (make_plan3(prob, 0, pl)) return 0;
  return (size_t)pl.nseg * 4 * (size_t)pl.S * pl.d.plane * sizeof(float);
}

int gbw_acoustic3d_forward(const gbw_acoustic3d_problem* prob, const gbw_acoustic3d_inputs* in,
                           const gbw_acoustic3d_outputs* out, void* workspace,
                           size_t workspace_bytes, void* stream) {
  Plan3 pl;
  int rc = make_plan3(prob, 0, pl);
  if (rc) return rc;
  if (!in || !out) return fail(GBW_ERR_INVALID, "inputs / outputs are NULL");
  if (!workspace || workspace_bytes < pl.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(pl.total) + " bytes");
  const bool rec = out->rec_p || out->rec_vx || out->rec_vy || out->rec_vz;
  if (rec && !(out->rec_p && out->rec_vx && out->rec_vy && out->rec_vz))
    return fail(GBW_ERR_INVALID, "rec_p, rec_vx, rec_vy and rec_vz must be given together");
  if (pl.k > 0 && !out->checkpoints)
    return fail(GBW_ERR_INVALID, "checkpoint_every > 0 needs a checkpoint buffer");
  const Dims3& d = pl.d;
  std::vector<int> src, rcell;
  if ((rc = geometry3(pl, in, src, rcell))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_src), src.data(), src.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_rcell), rcell.data(), rcell.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  // pageable copies are staged before hipMemcpyAsync returns, but keep the
  // host vectors alive until the copies are done regardless
  HIP_TRY(hipStreamSynchronize(st));
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
  // wavelet rows: shared (nt,), per shot (src_n, nt), per point (src_n, K, nt)
  a.wav = in->src_v;
  a.wav_stride = prob->wavelet_shared == 1 ? 0
                                           : (prob->wavelet_shared == 2 ? (long long)pl.K : 1) * prob->nt;
  a.wav_kstride = prob->wavelet_shared == 2 ? prob->nt : 0;
  a.src_k = pl.K;
  a.energy = energy ? E : nullptr;
  const dim3 grid(step_blocks3(d), pl.groups);
  float* cur = state;
  float* nxt = state + 4 * fstride;
  March3Args m{};
  m.d = d; m.S = pl.S; m.nt = pl.nt; m.dt = prob->dt; m.fstride = fstride; m.coef = coef;
  m.src_cell = a.src_cell; m.wav = a.wav; m.wav_stride = a.wav_stride; m.energy = a.energy;
  Fused3Args fa{};
  fa.d = d; fa.S = pl.S; fa.nt = pl.nt; fa.dt = prob->dt; fa.fstride = fstride; fa.coef = coef;
  fa.src_cell = a.src_cell; fa.wav = a.wav; fa.wav_stride = a.wav_stride; fa.energy = a.energy;
  if ((pl.march || pl.fused) && (rc = fill3(nxt, 4 * fstride, 0.f, st))) return rc;
  if (pl.fused) fused_setup3(pl, fa, prob->z_chunk);
  for (int it = 0; it < pl.nt; ++it) {
    if (pl.k > 0 && it % pl.k == 0)
      HIP_TRY(hipMemcpyAsync(static_cast<char*>(out->checkpoints) +
                                 (size_t)(it / pl.k) * 4 * fstride * sizeof(float),
                             cur, 4 * fstride * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (pl.fused) {
      fa.it = it; fa.A = cur; fa.B = nxt;
      if ((rc = fused_step3(pl, fa, st))) return rc;
      std::swap(cur, nxt);
    } else if (pl.march) {
      m.it = it; m.A = cur; m.B = nxt;
      if ((rc = march_step3(pl, m, false, st))) return rc;
      std::swap(cur, nxt);
    } else {
      a.it = it;
      if ((rc = forward_step3(a, grid, st))) return rc;
    }
    if (rec && pl.R > 0) {
      record3_kernel<<<(unsigned)(((long long)pl.S * pl.R + 255) / 256), 256, 0, st>>>(
          d, pl.S, pl.R, pl.nt, it, cur, fstride, at3<int>(workspace, pl.off_rcell), out->rec_p,
          out->rec_vx, out->rec_vy, out->rec_vz);
      LAUNCH_CHECK();
    }
    if (energy && (((it + 1) % pl.echunk == 0) || it + 1 == pl.nt)) {
      energy3_flush_kernel<<<(unsigned)((4 * inner + 255) / 256), 256, 0, st>>>(
          E, pl.groups, inner, out->energy_p, out->energy_vx, out->energy_vy, out->energy_vz);
      LAUNCH_CHECK();
    }
  }
  return flush_timing(st);
}

int gbw_acoustic3d_backward(const gbw_acoustic3d_problem* prob, const gbw_acoustic3d_inputs* in,
                            const gbw_acoustic3d_adjoint* adj, void* workspace,
                            size_t workspace_bytes, void* stream) {
  if (!adj || !in) return fail(GBW_ERR_INVALID, "inputs / adjoint are NULL");
  const int pass = adj->grad_rho ? 2 : 1;
  Plan3 pl;
  int rc = make_plan3(prob, pass, pl);
  if (rc) return rc;
  if (!workspace || workspace_bytes < pl.total)
    return fail(GBW_ERR_WORKSPACE, "workspace too small: need " + std::to_string(pl.total) + " bytes");
  if (!adj->checkpoints) return fail(GBW_ERR_INVALID, "checkpoints are NULL");
  const Dims3& d = pl.d;
  std::vector<int> src, rcell;
  if ((rc = geometry3(pl, in, src, rcell))) return rc;
  // receivers grouped by cell, first appearance order; receivers of a cell in their order
  std::vector<int> ucell, ustart, urcv;
  {
    std::vector<int> order(pl.R);
    for (int r = 0; r < pl.R; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rcell[a] < rcell[b]; });
    for (int i = 0; i < pl.R; ++i) {
      if (i == 0 || rcell[order[i]] != rcell[order[i - 1]]) {
        ucell.push_back(rcell[order[i]]);
        ustart.push_back(i);
      }
      urcv.push_back(order[i]);
    }
    ustart.push_back(pl.R);
    if (ucell.empty()) { ucell.push_back(0); urcv.push_back(0); }
  }
  // per (z, y) row of the plane, the range of its receiver cells in ucell
  std::vector<int> rowstart((size_t)d.NZ * d.NY + 1, 0);
  for (int u = 0; u + 1 < (int)ustart.size(); ++u) ++rowstart[ucell[u] / d.PX + 1];
  for (size_t i = 1; i < rowstart.size(); ++i) rowstart[i] += rowstart[i - 1];
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_src), src.data(), src.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_ucell), ucell.data(), ucell.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_ustart), ustart.data(),
                         ustart.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_urcv), urcv.data(), urcv.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(at3<int>(workspace, pl.off_rowstart), rowstart.data(),
                         rowstart.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));  // the host vectors go out of scope
  float* coef = at3<float>(workspace, pl.off_coef);
  coef3_kernel<<<(unsigned)((d.plane + 255) / 256), 256, 0, st>>>(d, in->vp, in->rho, in->damp,
                                                                    prob->dt, prob->dz, coef);
  LAUNCH_CHECK();
  const long long fstride = (long long)pl.S * d.plane;
  float* state = at3<float>(workspace, pl.off_state);
  float* lam = at3<float>(workspace, pl.off_lam);
  float* r = at3<float>(workspace, pl.off_r);
  float* D = at3<float>(workspace, pl.off_D);
  float* P1 = pass == 2 ? at3<float>(workspace, pl.off_P1) : nullptr;
  float* gA1 = at3<float>(workspace, pl.off_gA1);
  float* gA2 = pass == 2 ? at3<float>(workspace, pl.off_gA2) : nullptr;
  if ((rc = fill3(lam, 4 * fstride, 0.f, st))) return rc;
  if ((rc = fill3(r, fstride, 0.f, st))) return rc;
  if ((rc = fill3(gA1, (long long)pl.groups * d.plane, 0.f, st))) return rc;
  if (gA2 && (rc = fill3(gA2, (long long)pl.groups * d.plane, 0.f, st))) return rc;
  const bool shared = prob->wavelet_shared == 1;
  const long long nw = shared ? pl.nt : (long long)pl.S * (prob->wavelet_shared == 2 ? pl.K : 1) * pl.nt;
  if (adj->grad_src_v && (rc = fill3(adj->grad_src_v, nw, 0.f, st))) return rc;
  // several points per shot: per-point samples, reduced to the wavelet's shape at the end
  float* gwp = (pl.K > 1 && adj->grad_src_v) ? at3<float>(workspace, pl.off_gwp) : nullptr;
  if (gwp && (rc = fill3(gwp, (long long)pl.S * pl.K * pl.nt, 0.f, st))) return rc;
  Step3Args a{};
  a.d = d; a.S = pl.S; a.spb = pl.spb; a.nt = pl.nt; a.dt = prob->dt;
  a.state = state; a.fstride = fstride; a.coef = coef;
  a.src_cell = at3<int>(workspace, pl.off_src);
  a.wav = in->src_v;
  a.wav_stride = shared ? 0 : (prob->wavelet_shared == 2 ? (long long)pl.K : 1) * prob->nt;
  a.wav_kstride = prob->wavelet_shared == 2 ? prob->nt : 0;
  a.src_k = pl.K;
  March3Args m{};
  m.d = d; m.S = pl.S; m.nt = pl.nt; m.dt = prob->dt; m.fstride = fstride; m.coef = coef;
  m.src_cell = a.src_cell; m.wav = a.wav; m.wav_stride = a.wav_stride;
  Adj3Args b{};
  b.d = d; b.S = pl.S; b.spb = pl.spb; b.nt = pl.nt; b.dt = prob->dt;
  b.lam = lam; b.fstride = fstride; b.r = r; b.coef = coef;
  b.gA1 = gA1; b.gA2 = gA2; b.src_cell = a.src_cell;
  b.gw = gwp ? gwp : (adj->grad_src_v && !shared) ? adj->grad_src_v : nullptr;
  b.src_k = pl.K;
  b.gws = (adj->grad_src_v && shared) ? at3<float>(workspace, pl.off_gws) : nullptr;
  const dim3 grid(step_blocks3(d), pl.groups);
  if (pl.R > 0 && (adj->g_p || adj->g_vx || adj->g_vy || adj->g_vz)) {
    b.rowstart = at3<int>(workspace, pl.off_rowstart);
    b.ucell = at3<int>(workspace, pl.off_ucell);
    b.ustart = at3<int>(workspace, pl.off_ustart);
    b.urcv = at3<int>(workspace, pl.off_urcv);
    b.R = pl.R;
    b.g_p = adj->g_p; b.g_vx = adj->g_vx; b.g_vy = adj->g_vy; b.g_vz = adj->g_vz;
  }
  for (int c = pl.nseg - 1; c >= 0; --c) {
    const int t0 = c * pl.k, t1 = std::min(pl.nt, t0 + pl.k);
    HIP_TRY(hipMemcpyAsync(state,
                           static_cast<const char*>(adj->checkpoints) +
                               (size_t)c * 4 * fstride * sizeof(float),
                           4 * fstride * sizeof(float), hipMemcpyDeviceToDevice, st));
    float* cur = state;
    float* nxt = state + 4 * fstride;
    for (int it = t0; it < t1; ++it) {  // recompute: brackets (and p) of every step
      float* dst = D + (size_t)(it - t0) * fstride;
      float* pst = P1 ? P1 + (size_t)(it - t0) * fstride : nullptr;
      if (pl.march) {
        m.it = it; m.A = cur; m.B = nxt; m.dstore = dst; m.pstore = pst;
        if ((rc = march_step3(pl, m, true, st))) return rc;
        std::swap(cur, nxt);
      } else {
        a.it = it; a.dstore = dst; a.pstore = pst;
        if ((rc = forward_step3(a, grid, st))) return rc;
      }
    }
    for (int it = t1 - 1; it >= t0; --it) {
      b.it = it;
      b.vpart = it + 1 < pl.nt;  // r of step it+1 is in the buffer
      b.dstep = D + (size_t)(it - t0) * fstride;
      b.pstep = P1 ? P1 + (size_t)(it - t0) * fstride : nullptr;
      {
        LaunchTimer timer("adj3_vq_kernel", st);
        adj3_vq_kernel<<<grid, T3, 0, st>>>(b);
      }
      LAUNCH_CHECK();
      {
        LaunchTimer timer("adj3_p_kernel", st);
        if (b.src_k > 1)
          adj3_p_kernel<true><<<grid, T3, 0, st>>>(b);
        else
          adj3_p_kernel<false><<<grid, T3, 0, st>>>(b);
      }
      LAUNCH_CHECK();
      if (b.gws) {
        wavelet_sum3_kernel<<<1, 64, 0, st>>>(b.gws, pl.S, adj->grad_src_v, it);
        LAUNCH_CHECK();
      }
    }
  }
  if (gwp) {  // per-point samples -> (src_n, K, nt) as they are, or summed per shot
    if (prob->wavelet_shared == 2)
      HIP_TRY(hipMemcpyAsync(adj->grad_src_v, gwp, (size_t)nw * sizeof(float),
                             hipMemcpyDeviceToDevice, st));
    else {
      wavelet_points3_kernel<<<(unsigned)((nw + 255) / 256), 256, 0, st>>>(gwp, pl.S, pl.K, pl.nt,
                                                                         adj->grad_src_v);
      LAUNCH_CHECK();
    }
  }
  if (adj->grad_vp || adj->grad_rho) {
    const long long inner = (long long)d.nz * d.ny * d.nx;
    fold3_kernel<<<(unsigned)((inner + 255) / 256), 256, 0, st>>>(
        d, pl.groups, gA1, gA2, in->vp, in->rho, prob->dt, prob->dz, adj->grad_vp, adj->grad_rho);
    LAUNCH_CHECK();
  }
  return flush_timing(st);
}

}  // extern "C"
