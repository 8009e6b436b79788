// CPU oracle for the 3D acoustic staggered-grid propagator (test infrastructure).
//
// This file is the checker, not the product: only tests, the smoke check and
// the CPU-baseline leg of the benchmark load it.  It restates, in plain C++,
// GeoBrain's `acoustic_forward_kernel_3d`:
//   geobrain/physics/wave/kernels/acoustic_2d.py:70-120   pad_model_3d (edge replicate)
//   geobrain/physics/wave/kernels/acoustic_3d.py:401-431  c1, c2, alpha1, kappa1, alpha2,
//                                                         kappa2 (x), kappa3 (y), kappa4 (z)
//   geobrain/physics/wave/kernels/acoustic_3d.py:124-285  one time step (p, source, free
//                                                         surface, vx, vy, vz, free surface,
//                                                         receivers, energy)
//   geobrain/physics/wave/kernels/acoustic_3d.py:433-476  torch.chunk chunking of the energy sums
// and the exact discrete adjoint of that step (gradient w.r.t. vp, rho and the
// source wavelet), derived by transposing each sub-step, with forward
// checkpoints every `chunk` steps and recomputation in between.
//
// Two precisions from one template, as the 2D oracle: float rounds every
// operation to FP32 in the reference's op order (build with -ffp-contract=off);
// double runs the same arithmetic in FP64 from the same FP32 inputs.

#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

template <typename T>
static T dt_as(double v) { return (T)v; }

static inline int clamp3(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <typename T>
struct Grid3 {
  int nz, ny, nx, nabc, fs, NZ, NY, NX, f;
  T dt, dz;
  size_t cells() const { return (size_t)NZ * NY * NX; }
  size_t at(int i, int j, int k) const { return ((size_t)i * NY + j) * NX + k; }
};

template <typename T>
struct Coefs3 {
  std::vector<T> a1, omk1, a2, omk2, omk3, omk4;
};

template <typename T>
static void build_coefs3(const Grid3<T>& g, const float* vp, const float* rho, const float* damp,
                         Coefs3<T>& c) {
  const size_t n = g.cells();
  c.a1.resize(n); c.omk1.resize(n); c.a2.resize(n);
  c.omk2.resize(n); c.omk3.resize(n); c.omk4.resize(n);
  for (int i = 0; i < g.NZ; ++i) {
    const int a = clamp3(i - g.nabc, 0, g.nz - 1);
    for (int j = 0; j < g.NY; ++j) {
      const int b = clamp3(j - g.nabc, 0, g.ny - 1);
      for (int k = 0; k < g.NX; ++k) {
        const int e = clamp3(k - g.nabc, 0, g.nx - 1);
        const size_t q = g.at(i, j, k);
        const size_t m = ((size_t)a * g.ny + b) * g.nx + e;
        const T vc = (T)vp[m];
        const T den = (T)rho[m];
        // alpha1 = den * c * c * dt / dz   (acoustic_3d.py:403)
        T t = den * vc;
        t = t * vc;
        t = t * g.dt;
        c.a1[q] = t / g.dz;
        // kappa1 = damp * dt; the step uses (1.0 - kappa1)   (:404, :159)
        const T k1 = (T)damp[q] * g.dt;
        c.omk1[q] = (T)1 - k1;
        // alpha2 = dt / (den * dz), evaluated as reciprocal(den * dz) * dt   (:407)
        const T dd = den * g.dz;
        const T r = (T)1 / dd;
        c.a2[q] = r * g.dt;
        // kappa2[:, :, 1:NX-2] = 0.5 * (damp[x] + damp[x+1]) * dt   (:410-413)
        T k2 = 0;
        if (k >= 1 && k <= g.NX - 3) {
          T s = (T)damp[q] + (T)damp[g.at(i, j, k + 1)];
          s = (T)0.5 * s;
          k2 = s * g.dt;
        }
        c.omk2[q] = (T)1 - k2;
        // kappa3[:, 1:NY-2, :] = 0.5 * (damp[y] + damp[y+1]) * dt   (:416-419)
        T k3 = 0;
        if (j >= 1 && j <= g.NY - 3) {
          T s = (T)damp[q] + (T)damp[g.at(i, j + 1, k)];
          s = (T)0.5 * s;
          k3 = s * g.dt;
        }
        c.omk3[q] = (T)1 - k3;
        // kappa4[f:NZ-2, :, :] = 0.5 * (damp[z] + damp[z+1]) * dt   (:422-426)
        T k4 = 0;
        if (i >= g.f && i <= g.NZ - 3) {
          T s = (T)damp[q] + (T)damp[g.at(i + 1, j, k)];
          s = (T)0.5 * s;
          k4 = s * g.dt;
        }
        c.omk4[q] = (T)1 - k4;
      }
    }
  }
}

template <typename T>
static inline T c1v() { return (T)(9.0 / 8.0); }
template <typename T>
static inline T c2v() { return (T)(-1.0 / 24.0); }

// c1 * (a - b) + c2 * (c - d), one rounding per operation
template <typename T>
static inline T d4(T a, T b, T c, T d) {
  const T t1 = c1v<T>() * (a - b);
  const T t2 = c2v<T>() * (c - d);
  return t1 + t2;
}

template <typename T>
struct State3 {
  std::vector<T> p, vx, vy, vz;
  void init(size_t n) { p.assign(n, 0); vx.assign(n, 0); vy.assign(n, 0); vz.assign(n, 0); }
};

// One time step `it` (acoustic_3d.py:124-239).
template <typename T>
static void step3(const Grid3<T>& g, const Coefs3<T>& c, int S, const std::vector<int>& sz,
                  const std::vector<int>& sy, const std::vector<int>& sx, const float* w, int nt,
                  int it, State3<T>& st, int K = 1, int per_point = 0) {
  const size_t n = g.cells();
  const int NY = g.NY, NX = g.NX;
  const size_t sy1 = (size_t)NX, sz1 = (size_t)NY * NX;
  // p on z in [f+1, NZ-3], y in [2, NY-3], x in [2, NX-3]   (:129-162)
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < S; ++s)
    for (int i = g.f + 1; i <= g.NZ - 3; ++i) {
      T* p = st.p.data() + s * n;
      const T* vx = st.vx.data() + s * n;
      const T* vy = st.vy.data() + s * n;
      const T* vz = st.vz.data() + s * n;
      for (int j = 2; j <= NY - 3; ++j)
        for (int k = 2; k <= NX - 3; ++k) {
          const size_t q = g.at(i, j, k);
          const T dvx = d4(vx[q], vx[q - 1], vx[q + 1], vx[q - 2]);
          const T dvy = d4(vy[q], vy[q - sy1], vy[q + sy1], vy[q - 2 * sy1]);
          const T dvz = d4(vz[q], vz[q - sz1], vz[q + sz1], vz[q - 2 * sz1]);
          T sum = dvx + dvy;
          sum = sum + dvz;
          const T u = c.omk1[q] * p[q];
          p[q] = u - c.a1[q] * sum;
        }
    }
  for (int s = 0; s < S; ++s) {
    T* p = st.p.data() + s * n;
    // source (:165-176): p[s, sz, sy, sx] += dt * w[s, it] at each point of
    // the shot (point k at s*K+k; distinct cells); a row per shot or per point
    for (int k = 0; k < K; ++k) {
      const size_t row = per_point ? (size_t)s * K + k : (size_t)s;
      const T upd = g.dt * (T)w[row * nt + it];
      const size_t q = g.at(sz[s * K + k], sy[s * K + k], sx[s * K + k]);
      p[q] = p[q] + upd;
    }
    if (g.fs)  // p[:, f-1] = -p[:, f+1]   (:179-180)
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NX; ++k) p[g.at(g.f - 1, j, k)] = -p[g.at(g.f + 1, j, k)];
  }
  // vx: z in [f, NZ-2], y in [1, NY-2], x in [1, NX-3]   (:184-199)
  // vy: z in [f, NZ-2], y in [1, NY-3], x in [1, NX-2]   (:203-218)
  // vz: z in [f, NZ-3], y in [1, NY-2], x in [1, NX-2]   (:222-237)
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < S; ++s)
    for (int i = g.f; i <= g.NZ - 2; ++i) {
      const T* p = st.p.data() + s * n;
      T* vx = st.vx.data() + s * n;
      T* vy = st.vy.data() + s * n;
      T* vz = st.vz.data() + s * n;
      for (int j = 1; j <= NY - 2; ++j)
        for (int k = 1; k <= NX - 2; ++k) {
          const size_t q = g.at(i, j, k);
          if (k <= NX - 3) {
            const T u = c.omk2[q] * vx[q];
            vx[q] = u - c.a2[q] * d4(p[q + 1], p[q], p[q + 2], p[q - 1]);
          }
          if (j <= NY - 3) {
            const T u = c.omk3[q] * vy[q];
            vy[q] = u - c.a2[q] * d4(p[q + sy1], p[q], p[q + 2 * sy1], p[q - sy1]);
          }
          if (i <= g.NZ - 3) {
            const T u = c.omk4[q] * vz[q];
            vz[q] = u - c.a2[q] * d4(p[q + sz1], p[q], p[q + 2 * sz1], p[q - sz1]);
          }
        }
    }
  if (g.fs)  // vz[:, f-1] = vz[:, f]   (:240-241)
    for (int s = 0; s < S; ++s) {
      T* vz = st.vz.data() + s * n;
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NX; ++k) vz[g.at(g.f - 1, j, k)] = vz[g.at(g.f, j, k)];
    }
}

// acc[a, b, e] += sum_s f[s, a+nabc, b+nabc, e+nabc]^2   (:250-262)
template <typename T>
static void energy3(const Grid3<T>& g, int S, const std::vector<T>& fld, T* acc) {
  const size_t n = g.cells();
#pragma omp parallel for collapse(2) schedule(static)
  for (int a = 0; a < g.nz; ++a)
    for (int b = 0; b < g.ny; ++b)
      for (int e = 0; e < g.nx; ++e) {
        const size_t q = g.at(a + g.nabc, b + g.nabc, e + g.nabc);
        T sum = fld[q] * fld[q];
        for (int s = 1; s < S; ++s) {
          const T v = fld[s * n + q];
          sum = sum + v * v;
        }
        const size_t o = ((size_t)a * g.ny + b) * g.nx + e;
        acc[o] = acc[o] + sum;
      }
}

template <typename T>
static int forward3_impl(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                         int S, const int64_t* src_z, const int64_t* src_y, const int64_t* src_x,
                         const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_y,
                         const int64_t* rcv_x, const float* vp, const float* rho,
                         const float* damp, int chunk_len, T* rp, T* rvx, T* rvy, T* rvz, T* ep,
                         T* evx, T* evy, T* evz, T* final_state, int K = 1, int per_point = 0) {
  if (nz <= 0 || ny <= 0 || nx <= 0 || nabc <= 0 || nt <= 0 || S <= 0 || R < 0 || K <= 0)
    return -1;
  Grid3<T> g;
  g.nz = nz; g.ny = ny; g.nx = nx; g.nabc = nabc; g.fs = fs ? 1 : 0;
  g.NZ = nz + 2 * nabc; g.NY = ny + 2 * nabc; g.NX = nx + 2 * nabc;
  g.f = fs ? nabc : 1;
  g.dt = dt_as<T>(dt); g.dz = dt_as<T>(dz);
  std::vector<int> sz((size_t)S * K), sy((size_t)S * K), sx((size_t)S * K), rz(R), ry(R), rx(R);
  for (int s = 0; s < S * K; ++s) {
    sz[s] = (int)src_z[s] + nabc; sy[s] = (int)src_y[s] + nabc; sx[s] = (int)src_x[s] + nabc;
    if (sz[s] < 0 || sz[s] >= g.NZ || sy[s] < 0 || sy[s] >= g.NY || sx[s] < 0 || sx[s] >= g.NX)
      return -2;
  }
  for (int r = 0; r < R; ++r) {  // vz has NZ-1 planes, vy NY-1 rows, vx NX-1 columns
    rz[r] = (int)rcv_z[r] + nabc; ry[r] = (int)rcv_y[r] + nabc; rx[r] = (int)rcv_x[r] + nabc;
    if (rz[r] < 0 || rz[r] >= g.NZ - 1 || ry[r] < 0 || ry[r] >= g.NY - 1 || rx[r] < 0 ||
        rx[r] >= g.NX - 1)
      return -3;
  }
  Coefs3<T> c;
  build_coefs3(g, vp, rho, damp, c);
  State3<T> st;
  const size_t n = g.cells();
  st.init((size_t)S * n);
  const size_t inner = (size_t)nz * ny * nx;
  T* outs[4] = {ep, evx, evy, evz};
  for (T* o : outs)
    if (o) std::fill(o, o + inner, (T)0);
  std::vector<T> part[4];
  for (auto& v : part) v.assign(inner, (T)0);
  if (chunk_len <= 0) chunk_len = nt;
  for (int t0 = 0; t0 < nt; t0 += chunk_len) {
    const int t1 = std::min(nt, t0 + chunk_len);
    for (auto& v : part) std::fill(v.begin(), v.end(), (T)0);
    for (int it = t0; it < t1; ++it) {
      step3(g, c, S, sz, sy, sx, src_v, nt, it, st, K, per_point);
      for (int s = 0; s < S; ++s)  // receivers (:244-248)
        for (int r = 0; r < R; ++r) {
          const size_t q = s * n + g.at(rz[r], ry[r], rx[r]);
          const size_t o = ((size_t)s * nt + it) * R + r;
          if (rp) rp[o] = st.p[q];
          if (rvx) rvx[o] = st.vx[q];
          if (rvy) rvy[o] = st.vy[q];
          if (rvz) rvz[o] = st.vz[q];
        }
      energy3(g, S, st.p, part[0].data());
      energy3(g, S, st.vx, part[1].data());
      energy3(g, S, st.vy, part[2].data());
      energy3(g, S, st.vz, part[3].data());
    }
    for (int f = 0; f < 4; ++f)  // per-chunk sums added into the totals (:467-470)
      if (outs[f])
        for (size_t q = 0; q < inner; ++q) outs[f][q] = outs[f][q] + part[f][q];
  }
  if (final_state) {
    const size_t m = (size_t)S * n;
    std::memcpy(final_state, st.p.data(), m * sizeof(T));
    std::memcpy(final_state + m, st.vx.data(), m * sizeof(T));
    std::memcpy(final_state + 2 * m, st.vy.data(), m * sizeof(T));
    std::memcpy(final_state + 3 * m, st.vz.data(), m * sizeof(T));
  }
  return 0;
}

// bracket of the vx / vy / vz updates at cell q: c1 (p[+1] - p[0]) + c2 (p[+2] - p[-1])
template <typename T>
static inline T vbracket(const T* p, size_t q, size_t st) {
  return d4(p[q + st], p[q], p[q + 2 * st], p[q - st]);
}

// Adjoint of forward step `it` given the states before (st0) and after (st1)
// it.  lam holds the adjoint of state it+1 (records of step `it` injected) on
// entry and the adjoint of state `it` on exit.
template <typename T>
static void step3_adjoint(const Grid3<T>& g, const Coefs3<T>& c, int S, const std::vector<int>& sz,
                          const std::vector<int>& sy, const std::vector<int>& sx, int nt, int it,
                          const State3<T>& st0, const State3<T>& st1, State3<T>& lam,
                          std::vector<T>& gA1, std::vector<T>& gA2, T* gw, int shared,
                          std::vector<T>& scratch, int K = 1, int per_point = 0) {
  const size_t n = g.cells();
  const int NZ = g.NZ, NY = g.NY, NX = g.NX, f = g.f;
  const size_t sy1 = (size_t)NX, sz1 = (size_t)NY * NX;
  const T c1 = c1v<T>(), c2 = c2v<T>();
  std::vector<T> stepA1(n, (T)0), stepA2(n, (T)0);
  T* qx = scratch.data();
  T* qy = scratch.data() + n;
  T* qz = scratch.data() + 2 * n;
  for (int s = 0; s < S; ++s) {
    T* lp = lam.p.data() + s * n;
    T* lx = lam.vx.data() + s * n;
    T* ly = lam.vy.data() + s * n;
    T* lz = lam.vz.data() + s * n;
    const T* P1 = st1.p.data() + s * n;
    const T* VX0 = st0.vx.data() + s * n;
    const T* VY0 = st0.vy.data() + s * n;
    const T* VZ0 = st0.vz.data() + s * n;
    if (g.fs)  // transpose of vz[f-1] = vz[f]
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NX; ++k) {
          lz[g.at(f, j, k)] += lz[g.at(f - 1, j, k)];
          lz[g.at(f - 1, j, k)] = 0;
        }
    std::fill(qx, qx + 3 * n, (T)0);
    // transposes of the vz, vy, vx updates (reverse of their forward order):
    // q = alpha2 * lambda_v, d alpha2 -= lambda_v * bracket(p'), lambda_v *= 1 - kappa
#pragma omp parallel for schedule(static)
    for (int i = f; i <= NZ - 2; ++i)
      for (int j = 1; j <= NY - 2; ++j)
        for (int k = 1; k <= NX - 2; ++k) {
          const size_t q = g.at(i, j, k);
          if (i <= NZ - 3) {
            qz[q] = c.a2[q] * lz[q];
            stepA2[q] -= lz[q] * vbracket(P1, q, sz1);
            lz[q] = c.omk4[q] * lz[q];
          }
          if (j <= NY - 3) {
            qy[q] = c.a2[q] * ly[q];
            stepA2[q] -= ly[q] * vbracket(P1, q, sy1);
            ly[q] = c.omk3[q] * ly[q];
          }
          if (k <= NX - 3) {
            qx[q] = c.a2[q] * lx[q];
            stepA2[q] -= lx[q] * vbracket(P1, q, 1);
            lx[q] = c.omk2[q] * lx[q];
          }
        }
    // lambda_p' += -D^T q (gather form; q is zero outside the v regions)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NZ; ++i)
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NX; ++k) {
          auto QZ = [&](int a) -> T { return (a >= 0 && a < NZ) ? qz[g.at(a, j, k)] : (T)0; };
          auto QY = [&](int b) -> T { return (b >= 0 && b < NY) ? qy[g.at(i, b, k)] : (T)0; };
          auto QX = [&](int e) -> T { return (e >= 0 && e < NX) ? qx[g.at(i, j, e)] : (T)0; };
          T acc = 0;
          acc += c1 * (QZ(i) - QZ(i - 1)) + c2 * (QZ(i + 1) - QZ(i - 2));
          acc += c1 * (QY(j) - QY(j - 1)) + c2 * (QY(j + 1) - QY(j - 2));
          acc += c1 * (QX(k) - QX(k - 1)) + c2 * (QX(k + 1) - QX(k - 2));
          lp[g.at(i, j, k)] += acc;
        }
    if (g.fs)  // transpose of p[f-1] = -p[f+1]
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NX; ++k) {
          lp[g.at(f + 1, j, k)] -= lp[g.at(f - 1, j, k)];
          lp[g.at(f - 1, j, k)] = 0;
        }
    if (gw)  // transpose of the source terms
      for (int k = 0; k < K; ++k) {
        const T v = g.dt * lp[g.at(sz[s * K + k], sy[s * K + k], sx[s * K + k])];
        if (shared) gw[it] += v;
        else gw[(per_point ? (size_t)s * K + k : (size_t)s) * nt + it] += v;
      }
    // transpose of the p update: r = alpha1 * lambda_p', d alpha1 -= lambda_p' * bracket(v)
    T* r = qz;
    std::fill(r, r + n, (T)0);
#pragma omp parallel for schedule(static)
    for (int i = f + 1; i <= NZ - 3; ++i)
      for (int j = 2; j <= NY - 3; ++j)
        for (int k = 2; k <= NX - 3; ++k) {
          const size_t q = g.at(i, j, k);
          const T dvx = d4(VX0[q], VX0[q - 1], VX0[q + 1], VX0[q - 2]);
          const T dvy = d4(VY0[q], VY0[q - sy1], VY0[q + sy1], VY0[q - 2 * sy1]);
          const T dvz = d4(VZ0[q], VZ0[q - sz1], VZ0[q + sz1], VZ0[q - 2 * sz1]);
          T sum = dvx + dvy;
          sum = sum + dvz;
          r[q] = c.a1[q] * lp[q];
          stepA1[q] -= lp[q] * sum;
          lp[q] = c.omk1[q] * lp[q];
        }
    // lambda_v += -B^T r, B the divergence brackets (vx: NX-1 columns, vy: NY-1 rows, vz: NZ-1 planes)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NZ; ++i)
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NX; ++k) {
          auto RR = [&](int a, int b, int e) -> T {
            return (a >= 0 && a < NZ && b >= 0 && b < NY && e >= 0 && e < NX) ? r[g.at(a, b, e)]
                                                                               : (T)0;
          };
          const size_t q = g.at(i, j, k);
          if (k <= NX - 2)
            lx[q] += c1 * (RR(i, j, k + 1) - RR(i, j, k)) + c2 * (RR(i, j, k + 2) - RR(i, j, k - 1));
          if (j <= NY - 2)
            ly[q] += c1 * (RR(i, j + 1, k) - RR(i, j, k)) + c2 * (RR(i, j + 2, k) - RR(i, j - 1, k));
          if (i <= NZ - 2)
            lz[q] += c1 * (RR(i + 1, j, k) - RR(i, j, k)) + c2 * (RR(i + 2, j, k) - RR(i - 1, j, k));
        }
  }
  for (size_t q = 0; q < n; ++q) {
    gA1[q] += stepA1[q];
    gA2[q] += stepA2[q];
  }
}

template <typename T>
static int gradient3_impl(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                          int S, const int64_t* src_z, const int64_t* src_y, const int64_t* src_x,
                          const float* src_v, int shared, int R, const int64_t* rcv_z,
                          const int64_t* rcv_y, const int64_t* rcv_x, const float* vp,
                          const float* rho, const float* damp, int chunk_len, const T* g_p,
                          const T* g_vx, const T* g_vy, const T* g_vz, T* grad_vp, T* grad_rho,
                          T* grad_w, int K = 1, int per_point = 0) {
  if (nz <= 0 || ny <= 0 || nx <= 0 || nabc <= 0 || nt <= 0 || S <= 0 || R < 0 || K <= 0)
    return -1;
  Grid3<T> g;
  g.nz = nz; g.ny = ny; g.nx = nx; g.nabc = nabc; g.fs = fs ? 1 : 0;
  g.NZ = nz + 2 * nabc; g.NY = ny + 2 * nabc; g.NX = nx + 2 * nabc;
  g.f = fs ? nabc : 1;
  g.dt = dt_as<T>(dt); g.dz = dt_as<T>(dz);
  std::vector<int> sz((size_t)S * K), sy((size_t)S * K), sx((size_t)S * K), rz(R), ry(R), rx(R);
  for (int s = 0; s < S * K; ++s) {
    sz[s] = (int)src_z[s] + nabc; sy[s] = (int)src_y[s] + nabc; sx[s] = (int)src_x[s] + nabc;
    if (sz[s] < 0 || sz[s] >= g.NZ || sy[s] < 0 || sy[s] >= g.NY || sx[s] < 0 || sx[s] >= g.NX)
      return -2;
  }
  for (int r = 0; r < R; ++r) {
    rz[r] = (int)rcv_z[r] + nabc; ry[r] = (int)rcv_y[r] + nabc; rx[r] = (int)rcv_x[r] + nabc;
    if (rz[r] < 0 || rz[r] >= g.NZ - 1 || ry[r] < 0 || ry[r] >= g.NY - 1 || rx[r] < 0 ||
        rx[r] >= g.NX - 1)
      return -3;
  }
  // the forward reads src_v as (S, nt); a shared wavelet is broadcast
  // (rows per point: S*K of them)
  const int rows = per_point ? S * K : S;
  std::vector<float> wfull((size_t)rows * nt);
  for (int s = 0; s < rows; ++s)
    for (int t = 0; t < nt; ++t) wfull[(size_t)s * nt + t] = src_v[shared ? t : (size_t)s * nt + t];
  Coefs3<T> c;
  build_coefs3(g, vp, rho, damp, c);
  const size_t n = g.cells();
  if (chunk_len <= 0 || chunk_len > nt) chunk_len = nt;
  std::vector<int> bounds;
  for (int t = 0; t < nt; t += chunk_len) bounds.push_back(t);
  bounds.push_back(nt);
  const int nseg = (int)bounds.size() - 1;
  std::vector<State3<T>> ckpt(nseg);
  State3<T> st;
  st.init((size_t)S * n);
  for (int k = 0; k < nseg; ++k) {
    ckpt[k] = st;
    for (int it = bounds[k]; it < bounds[k + 1]; ++it)
      step3(g, c, S, sz, sy, sx, wfull.data(), nt, it, st, K, per_point);
  }
  State3<T> lam;
  lam.init((size_t)S * n);
  std::vector<T> gA1(n, (T)0), gA2(n, (T)0), scratch(3 * n);
  if (grad_w) std::fill(grad_w, grad_w + (shared ? (size_t)nt : (size_t)rows * nt), (T)0);
  for (int k = nseg - 1; k >= 0; --k) {
    const int t0 = bounds[k], t1 = bounds[k + 1];
    std::vector<State3<T>> seg(t1 - t0 + 1);
    seg[0] = ckpt[k];
    for (int it = t0; it < t1; ++it) {
      seg[it - t0 + 1] = seg[it - t0];
      step3(g, c, S, sz, sy, sx, wfull.data(), nt, it, seg[it - t0 + 1], K, per_point);
    }
    for (int it = t1 - 1; it >= t0; --it) {
      for (int s = 0; s < S; ++s)  // records of step `it` read state it+1
        for (int r = 0; r < R; ++r) {
          const size_t q = s * n + g.at(rz[r], ry[r], rx[r]);
          const size_t o = ((size_t)s * nt + it) * R + r;
          if (g_p) lam.p[q] += g_p[o];
          if (g_vx) lam.vx[q] += g_vx[o];
          if (g_vy) lam.vy[q] += g_vy[o];
          if (g_vz) lam.vz[q] += g_vz[o];
        }
      step3_adjoint(g, c, S, sz, sy, sx, nt, it, seg[it - t0], seg[it - t0 + 1], lam, gA1, gA2,
                    grad_w, shared, scratch, K, per_point);
    }
  }
  // chain rule: alpha1 = den c c dt / dz, alpha2 = dt / (den dz), then fold pad_model_3d
  if (grad_vp) std::fill(grad_vp, grad_vp + (size_t)nz * ny * nx, (T)0);
  if (grad_rho) std::fill(grad_rho, grad_rho + (size_t)nz * ny * nx, (T)0);
  for (int i = 0; i < g.NZ; ++i) {
    const int a = clamp3(i - nabc, 0, nz - 1);
    for (int j = 0; j < g.NY; ++j) {
      const int b = clamp3(j - nabc, 0, ny - 1);
      for (int k = 0; k < g.NX; ++k) {
        const int e = clamp3(k - nabc, 0, nx - 1);
        const size_t q = g.at(i, j, k);
        const size_t m = ((size_t)a * ny + b) * nx + e;
        const T vc = (T)vp[m], den = (T)rho[m];
        if (grad_vp) grad_vp[m] += gA1[q] * ((T)2 * den * vc * g.dt / g.dz);
        if (grad_rho)
          grad_rho[m] += gA1[q] * (vc * vc * g.dt / g.dz) - gA2[q] * (g.dt / (den * den * g.dz));
      }
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int oracle3_forward_f32(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                        int S, const int64_t* src_z, const int64_t* src_y, const int64_t* src_x,
                        const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_y,
                        const int64_t* rcv_x, const float* vp, const float* rho,
                        const float* damp, int chunk_len, float* rp, float* rvx, float* rvy,
                        float* rvz, float* ep, float* evx, float* evy, float* evz,
                        float* final_state) {
  return forward3_impl<float>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v, R,
                              rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, rp, rvx, rvy, rvz,
                              ep, evx, evy, evz, final_state);
}

int oracle3_forward_f64(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                        int S, const int64_t* src_z, const int64_t* src_y, const int64_t* src_x,
                        const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_y,
                        const int64_t* rcv_x, const float* vp, const float* rho,
                        const float* damp, int chunk_len, double* rp, double* rvx, double* rvy,
                        double* rvz, double* ep, double* evx, double* evy, double* evz,
                        double* final_state) {
  return forward3_impl<double>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v,
                               R, rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, rp, rvx, rvy,
                               rvz, ep, evx, evy, evz, final_state);
}

int oracle3_gradient_f32(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                         int S, const int64_t* src_z, const int64_t* src_y, const int64_t* src_x,
                         const float* src_v, int shared, int R, const int64_t* rcv_z,
                         const int64_t* rcv_y, const int64_t* rcv_x, const float* vp,
                         const float* rho, const float* damp, int chunk_len, const float* g_p,
                         const float* g_vx, const float* g_vy, const float* g_vz, float* grad_vp,
                         float* grad_rho, float* grad_w) {
  return gradient3_impl<float>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v,
                               shared, R, rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, g_p, g_vx,
                               g_vy, g_vz, grad_vp, grad_rho, grad_w);
}

int oracle3_gradient_f64(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                         int S, const int64_t* src_z, const int64_t* src_y, const int64_t* src_x,
                         const float* src_v, int shared, int R, const int64_t* rcv_z,
                         const int64_t* rcv_y, const int64_t* rcv_x, const float* vp,
                         const float* rho, const float* damp, int chunk_len, const double* g_p,
                         const double* g_vx, const double* g_vy, const double* g_vz,
                         double* grad_vp, double* grad_rho, double* grad_w) {
  return gradient3_impl<double>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v,
                                shared, R, rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, g_p,
                                g_vx, g_vy, g_vz, grad_vp, grad_rho, grad_w);
}

// several source points per shot: src_* hold S*K points (point k of shot s
// at s*K+k), src_v a row per shot (per_point 0) or per point (1)
int oracle3_forward_mp_f32(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                           int S, int K, int per_point, const int64_t* src_z,
                           const int64_t* src_y, const int64_t* src_x, const float* src_v, int R,
                           const int64_t* rcv_z, const int64_t* rcv_y, const int64_t* rcv_x,
                           const float* vp, const float* rho, const float* damp, int chunk_len,
                           float* rp, float* rvx, float* rvy, float* rvz, float* ep, float* evx,
                           float* evy, float* evz) {
  return forward3_impl<float>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v, R,
                              rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, rp, rvx, rvy, rvz,
                              ep, evx, evy, evz, nullptr, K, per_point);
}

int oracle3_forward_mp_f64(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                           int S, int K, int per_point, const int64_t* src_z,
                           const int64_t* src_y, const int64_t* src_x, const float* src_v, int R,
                           const int64_t* rcv_z, const int64_t* rcv_y, const int64_t* rcv_x,
                           const float* vp, const float* rho, const float* damp, int chunk_len,
                           double* rp, double* rvx, double* rvy, double* rvz, double* ep,
                           double* evx, double* evy, double* evz) {
  return forward3_impl<double>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v,
                               R, rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, rp, rvx, rvy,
                               rvz, ep, evx, evy, evz, nullptr, K, per_point);
}

int oracle3_gradient_mp_f32(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                            int S, int K, int per_point, const int64_t* src_z,
                            const int64_t* src_y, const int64_t* src_x, const float* src_v, int R,
                            const int64_t* rcv_z, const int64_t* rcv_y, const int64_t* rcv_x,
                            const float* vp, const float* rho, const float* damp, int chunk_len,
                            const float* g_p, const float* g_vx, const float* g_vy,
                            const float* g_vz, float* grad_vp, float* grad_rho, float* grad_w) {
  return gradient3_impl<float>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v, 0,
                               R, rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, g_p, g_vx, g_vy,
                               g_vz, grad_vp, grad_rho, grad_w, K, per_point);
}

int oracle3_gradient_mp_f64(int nz, int ny, int nx, int nabc, int fs, double dt, double dz, int nt,
                            int S, int K, int per_point, const int64_t* src_z,
                            const int64_t* src_y, const int64_t* src_x, const float* src_v, int R,
                            const int64_t* rcv_z, const int64_t* rcv_y, const int64_t* rcv_x,
                            const float* vp, const float* rho, const float* damp, int chunk_len,
                            const double* g_p, const double* g_vx, const double* g_vy,
                            const double* g_vz, double* grad_vp, double* grad_rho,
                            double* grad_w) {
  return gradient3_impl<double>(nz, ny, nx, nabc, fs, dt, dz, nt, S, src_z, src_y, src_x, src_v, 0,
                                R, rcv_z, rcv_y, rcv_x, vp, rho, damp, chunk_len, g_p, g_vx, g_vy,
                                g_vz, grad_vp, grad_rho, grad_w, K, per_point);
}

}  // extern "C"
