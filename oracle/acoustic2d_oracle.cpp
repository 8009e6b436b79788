// CPU oracle for the 2D acoustic staggered-grid propagator (test infrastructure).
//
// This file is the checker, not the product: only tests, the smoke check and
// the CPU-baseline leg of the benchmark load it.  It restates, in plain C++,
// the algorithm of GeoBrain's `acoustic_forward_kernel`:
//   geobrain/physics/wave/kernels/acoustic_2d.py:18-66    pad_model (edge replicate)
//   geobrain/physics/wave/kernels/acoustic_2d.py:402-420  c1, c2, alpha1, kappa1, alpha2, kappa2, kappa3
//   geobrain/physics/wave/kernels/acoustic_2d.py:211-285  one time step (p, source, free surface, vx, vz,
//                                                         receivers, energy)
//   geobrain/physics/wave/kernels/acoustic_2d.py:422-489  chunking (torch.chunk / snapshot chunks)
// and the exact discrete adjoint of that step (gradient w.r.t. vp, rho and the
// source wavelet), derived by transposing each sub-step.
//
// Two precisions are built from one template:
//   * float  : every operation rounded to FP32 in the reference's own op order
//              (no FMA contraction; build with -ffp-contract=off), so the forward
//              outputs reproduce the reference's FP32 numbers;
//   * double : the same arithmetic in FP64 from the same FP32 inputs, dt and dz
//              kept in double.  This is the FP64 oracle that errors are stated
//              against.
// The gradient uses forward checkpoints every `chunk` steps and recomputes the
// states of each chunk, so memory is O(chunk * state).

#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <cmath>

namespace {

template <typename T>
struct Grid {
  int nz, nx, nabc, fs, NZ, NX, f;
  T dt, dz;
  size_t cells() const { return (size_t)NZ * NX; }
  size_t at(int i, int j) const { return (size_t)i * NX + j; }
};

// Coefficient arrays on the padded grid, in the reference's op order.
template <typename T>
struct Coefs {
  std::vector<T> c, den, a1, omk1, a2, omk2, omk3;
};

template <typename T>
static T to_T_dt(double v);
template <>
float to_T_dt<float>(double v) { return (float)v; }
template <>
double to_T_dt<double>(double v) { return v; }

// pad_model (acoustic_2d.py:18-66): c[i,j] = v[clamp(i-nabc), clamp(j-nabc)].
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <typename T>
static void build_coefs(const Grid<T>& g, const float* vp, const float* rho, const float* damp,
                        Coefs<T>& k) {
  const size_t n = g.cells();
  k.c.resize(n); k.den.resize(n); k.a1.resize(n); k.omk1.resize(n);
  k.a2.resize(n); k.omk2.resize(n); k.omk3.resize(n);
  for (int i = 0; i < g.NZ; ++i) {
    const int a = clampi(i - g.nabc, 0, g.nz - 1);
    for (int j = 0; j < g.NX; ++j) {
      const int b = clampi(j - g.nabc, 0, g.nx - 1);
      const size_t q = g.at(i, j);
      const T c = (T)vp[(size_t)a * g.nx + b];
      const T den = (T)rho[(size_t)a * g.nx + b];
      k.c[q] = c;
      k.den[q] = den;
      // alpha1 = den * c * c * dt / dz        (acoustic_2d.py:407)
      T t = den * c;
      t = t * c;
      t = t * g.dt;
      k.a1[q] = t / g.dz;
      // kappa1 = damp * dt; the step uses (1.0 - kappa1)   (:408, :214)
      const T k1 = (T)damp[q] * g.dt;
      k.omk1[q] = (T)1 - k1;
      // alpha2 = dt / (den * dz), evaluated as reciprocal(den*dz) * dt   (:410)
      const T dd = den * g.dz;
      const T r = (T)1 / dd;
      k.a2[q] = r * g.dt;
      // kappa2[:, 1:NX-2] = 0.5 * (damp[:, j] + damp[:, j+1]) * dt   (:411-414)
      T k2 = 0;
      if (j >= 1 && j <= g.NX - 3) {
        T s = (T)damp[q] + (T)damp[g.at(i, j + 1)];
        s = (T)0.5 * s;
        k2 = s * g.dt;
      }
      k.omk2[q] = (T)1 - k2;
      // kappa3[f:NZ-2, :] = 0.5 * (damp[i, :] + damp[i+1, :]) * dt   (:416-420)
      T k3 = 0;
      if (i >= g.f && i <= g.NZ - 3) {
        T s = (T)damp[q] + (T)damp[g.at(i + 1, j)];
        s = (T)0.5 * s;
        k3 = s * g.dt;
      }
      k.omk3[q] = (T)1 - k3;
    }
  }
}

template <typename T>
struct State {
  std::vector<T> p, vx, vz;  // each S * NZ * NX (vx column NX-1 and vz row NZ-1 unused, kept 0)
  void init(size_t n) { p.assign(n, 0); vx.assign(n, 0); vz.assign(n, 0); }
};

template <typename T>
struct Geometry {
  int S, R, nt;
  int K = 1;                        // source points per shot
  int per_point = 0;                // 1: a wavelet row per point (S*K rows), else per shot
  std::vector<int> sz, sx, rz, rx;  // padded indices (sources: S*K, point k of shot s at s*K+k)
  const float* w;                   // wavelet rows
  int wrow(int s, int k) const { return per_point ? s * K + k : s; }
};

template <typename T>
static const T C1() { return (T)(9.0 / 8.0); }
template <typename T>
static const T C2() { return (T)(-1.0 / 24.0); }

// The bracket of the p-update (acoustic_2d.py:216-225), op order kept:
//   c1 * (((vx[j]-vx[j-1]) + vz[i]) - vz[i-1]) + c2 * (((vx[j+1]-vx[j-2]) + vz[i+1]) - vz[i-2])
template <typename T>
static inline T p_bracket(const Grid<T>& g, const T* vx, const T* vz, int i, int j) {
  T s1 = vx[g.at(i, j)] - vx[g.at(i, j - 1)];
  s1 = s1 + vz[g.at(i, j)];
  s1 = s1 - vz[g.at(i - 1, j)];
  T s2 = vx[g.at(i, j + 1)] - vx[g.at(i, j - 2)];
  s2 = s2 + vz[g.at(i + 1, j)];
  s2 = s2 - vz[g.at(i - 2, j)];
  const T t1 = C1<T>() * s1;
  const T t2 = C2<T>() * s2;
  return t1 + t2;
}

// vx bracket (:247-252): c1*(p[j+1]-p[j]) + c2*(p[j+2]-p[j-1])
template <typename T>
static inline T vx_bracket(const Grid<T>& g, const T* p, int i, int j) {
  const T d1 = p[g.at(i, j + 1)] - p[g.at(i, j)];
  const T d2 = p[g.at(i, j + 2)] - p[g.at(i, j - 1)];
  return C1<T>() * d1 + C2<T>() * d2;
}

// vz bracket (:259-264): c1*(p[i+1]-p[i]) + c2*(p[i+2]-p[i-1])
template <typename T>
static inline T vz_bracket(const Grid<T>& g, const T* p, int i, int j) {
  const T d1 = p[g.at(i + 1, j)] - p[g.at(i, j)];
  const T d2 = p[g.at(i + 2, j)] - p[g.at(i - 1, j)];
  return C1<T>() * d1 + C2<T>() * d2;
}

// One forward time step `it` for all shots (acoustic_2d.py:212-269).
template <typename T>
static void step_forward(const Grid<T>& g, const Coefs<T>& k, const Geometry<T>& geo, int it,
                         State<T>& st) {
  const size_t plane = g.cells();
  const int S = geo.S;
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < S; ++s) {
    for (int i = g.f + 1; i <= g.NZ - 3; ++i) {
      T* p = st.p.data() + s * plane;
      const T* vx = st.vx.data() + s * plane;
      const T* vz = st.vz.data() + s * plane;
      for (int j = 2; j <= g.NX - 3; ++j) {
        const size_t q = g.at(i, j);
        const T t1 = k.omk1[q] * p[q];
        const T br = p_bracket(g, vx, vz, i, j);
        p[q] = t1 - k.a1[q] * br;
      }
    }
  }
  for (int s = 0; s < S; ++s) {
    T* p = st.p.data() + s * plane;
    // source (:228-237): p[s, sz, sx] = p[s, sz, sx] + dt * w[s, it] at each
    // point of the shot (distinct cells: every point reads the updated field)
    for (int k = 0; k < geo.K; ++k) {
      const T upd = g.dt * (T)geo.w[(size_t)geo.wrow(s, k) * geo.nt + it];
      const size_t q = g.at(geo.sz[s * geo.K + k], geo.sx[s * geo.K + k]);
      p[q] = p[q] + upd;
    }
    // free surface for p (:239-241)
    if (g.fs)
      for (int j = 0; j < g.NX; ++j) p[g.at(g.f - 1, j)] = -p[g.at(g.f + 1, j)];
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int s = 0; s < S; ++s) {
    for (int i = g.f; i <= g.NZ - 2; ++i) {
      const T* p = st.p.data() + s * plane;
      T* vx = st.vx.data() + s * plane;
      T* vz = st.vz.data() + s * plane;
      for (int j = 1; j <= g.NX - 3; ++j) {  // vx region (:244)
        const size_t q = g.at(i, j);
        const T t1 = k.omk2[q] * vx[q];
        vx[q] = t1 - k.a2[q] * vx_bracket(g, p, i, j);
      }
      if (i <= g.NZ - 3) {
        for (int j = 1; j <= g.NX - 2; ++j) {  // vz region (:256)
          const size_t q = g.at(i, j);
          const T t1 = k.omk3[q] * vz[q];
          vz[q] = t1 - k.a2[q] * vz_bracket(g, p, i, j);
        }
      }
    }
  }
  if (g.fs) {  // (:267-269)
    for (int s = 0; s < S; ++s) {
      T* vz = st.vz.data() + s * plane;
      for (int j = 0; j < g.NX; ++j) vz[g.at(g.f - 1, j)] = vz[g.at(g.f, j)];
    }
  }
}

template <typename T>
static void record(const Grid<T>& g, const Geometry<T>& geo, int it, const State<T>& st, T* rp,
                   T* rvx, T* rvz) {
  const size_t plane = g.cells();
  for (int s = 0; s < geo.S; ++s)
    for (int r = 0; r < geo.R; ++r) {
      const size_t q = s * plane + g.at(geo.rz[r], geo.rx[r]);
      const size_t o = ((size_t)s * geo.nt + it) * geo.R + r;
      if (rp) rp[o] = st.p[q];
      if (rvx) rvx[o] = st.vx[q];
      if (rvz) rvz[o] = st.vz[q];
    }
}

// energy (:276-285): acc[i-nabc, j-nabc] += sum_s f[s,i,j]^2 over the interior
template <typename T>
static void energy(const Grid<T>& g, int S, const std::vector<T>& fld, T* acc) {
  const size_t plane = g.cells();
#pragma omp parallel for schedule(static)
  for (int a = 0; a < g.nz; ++a)
    for (int b = 0; b < g.nx; ++b) {
      const size_t q = g.at(a + g.nabc, b + g.nabc);
      T sum = fld[q] * fld[q];
      for (int s = 1; s < S; ++s) {
        const T v = fld[s * plane + q];
        sum = sum + v * v;
      }
      acc[(size_t)a * g.nx + b] = acc[(size_t)a * g.nx + b] + sum;
    }
}

template <typename T>
static int setup(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                 const int64_t* src_z, const int64_t* src_x, const float* src_v, int R,
                 const int64_t* rcv_z, const int64_t* rcv_x, Grid<T>& g, Geometry<T>& geo,
                 int K = 1, int per_point = 0) {
  if (nz <= 0 || nx <= 0 || nabc <= 0 || nt <= 0 || S <= 0 || R < 0 || K <= 0) return -1;
  g.nz = nz; g.nx = nx; g.nabc = nabc; g.fs = fs ? 1 : 0;
  g.NZ = nz + 2 * nabc; g.NX = nx + 2 * nabc;
  g.f = fs ? nabc : 1;
  g.dt = to_T_dt<T>(dt); g.dz = to_T_dt<T>(dz);
  geo.S = S; geo.R = R; geo.nt = nt; geo.w = src_v; geo.K = K; geo.per_point = per_point;
  geo.sz.resize((size_t)S * K); geo.sx.resize((size_t)S * K); geo.rz.resize(R); geo.rx.resize(R);
  for (int s = 0; s < S * K; ++s) {
    geo.sz[s] = (int)src_z[s] + nabc; geo.sx[s] = (int)src_x[s] + nabc;
    if (geo.sz[s] < 0 || geo.sz[s] >= g.NZ || geo.sx[s] < 0 || geo.sx[s] >= g.NX) return -2;
  }
  for (int r = 0; r < R; ++r) {
    geo.rz[r] = (int)rcv_z[r] + nabc; geo.rx[r] = (int)rcv_x[r] + nabc;
    if (geo.rz[r] < 0 || geo.rz[r] >= g.NZ - 1 || geo.rx[r] < 0 || geo.rx[r] >= g.NX - 1) return -3;
  }
  return 0;
}

// chunk boundaries: chunk_len > 0 -> [0,c), [c,2c), ... (torch.chunk with size c, or snapshot chunks)
static std::vector<int> chunk_starts(int nt, int chunk_len) {
  std::vector<int> b;
  if (chunk_len <= 0) chunk_len = nt;
  for (int t = 0; t < nt; t += chunk_len) b.push_back(t);
  b.push_back(nt);
  return b;
}

template <typename T>
static int forward_impl(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                        const int64_t* src_z, const int64_t* src_x, const float* src_v, int R,
                        const int64_t* rcv_z, const int64_t* rcv_x, const float* vp,
                        const float* rho, const float* damp, int chunk_len, T* rp, T* rvx,
                        T* rvz, T* ep, T* evx, T* evz, T* snapshots, T* final_state,
                        int K = 1, int per_point = 0) {
  Grid<T> g; Geometry<T> geo;
  int rc = setup<T>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, R, rcv_z, rcv_x, g, geo,
                    K, per_point);
  if (rc) return rc;
  Coefs<T> k; build_coefs(g, vp, rho, damp, k);
  State<T> st; st.init((size_t)S * g.cells());
  const size_t inner = (size_t)nz * nx;
  std::vector<T> cp(inner), cvx(inner), cvz(inner);
  if (ep) std::fill(ep, ep + inner, (T)0);
  if (evx) std::fill(evx, evx + inner, (T)0);
  if (evz) std::fill(evz, evz + inner, (T)0);
  std::vector<int> bounds = chunk_starts(nt, chunk_len);
  for (size_t c = 0; c + 1 < bounds.size(); ++c) {
    std::fill(cp.begin(), cp.end(), (T)0);
    std::fill(cvx.begin(), cvx.end(), (T)0);
    std::fill(cvz.begin(), cvz.end(), (T)0);
    for (int it = bounds[c]; it < bounds[c + 1]; ++it) {
      step_forward(g, k, geo, it, st);
      record(g, geo, it, st, rp, rvx, rvz);
      energy(g, S, st.p, cp.data());
      energy(g, S, st.vx, cvx.data());
      energy(g, S, st.vz, cvz.data());
    }
    // per-chunk partial sums added into the totals (acoustic_2d.py:479-481)
    for (size_t q = 0; q < inner; ++q) {
      if (ep) ep[q] = ep[q] + cp[q];
      if (evx) evx[q] = evx[q] + cvx[q];
      if (evz) evz[q] = evz[q] + cvz[q];
    }
    if (snapshots) {  // shot 0 pressure interior after each chunk (:483-487)
      for (int a = 0; a < nz; ++a)
        for (int b = 0; b < nx; ++b)
          snapshots[c * inner + (size_t)a * nx + b] = st.p[g.at(a + nabc, b + nabc)];
    }
  }
  if (final_state) {
    const size_t n = (size_t)S * g.cells();
    std::memcpy(final_state, st.p.data(), n * sizeof(T));
    std::memcpy(final_state + n, st.vx.data(), n * sizeof(T));
    std::memcpy(final_state + 2 * n, st.vz.data(), n * sizeof(T));
  }
  return 0;
}

// Adjoint of one forward step `it`, given the forward states before (st0) and
// after (st1) the step.  lam holds the adjoint of state it+1 (records of step
// `it` already injected) on entry and the adjoint of state `it` on exit.
template <typename T>
static void step_adjoint(const Grid<T>& g, const Coefs<T>& k, const Geometry<T>& geo, int it,
                         const State<T>& st0, const State<T>& st1, State<T>& lam,
                         std::vector<T>& gA1, std::vector<T>& gA2, T* gw, int wavelet_shared,
                         std::vector<T>& scratch_q) {
  const size_t plane = g.cells();
  const int S = geo.S;
  const int NZ = g.NZ, NX = g.NX, f = g.f;
  const T c1 = C1<T>(), c2 = C2<T>();
  // per-step sums over shots (mirrors autograd's reduction of a broadcast coefficient)
  std::vector<T> stepA1(plane, (T)0), stepA2(plane, (T)0);
  for (int s = 0; s < S; ++s) {
    T* lp = lam.p.data() + s * plane;
    T* lx = lam.vx.data() + s * plane;
    T* lz = lam.vz.data() + s * plane;
    const T* P1 = st1.p.data() + s * plane;  // p after the step (p' after source + mirror)
    const T* VX0 = st0.vx.data() + s * plane;
    const T* VZ0 = st0.vz.data() + s * plane;
    // 6^T: vz[f-1] = vz[f]
    if (g.fs)
      for (int j = 0; j < NX; ++j) {
        lz[g.at(f, j)] += lz[g.at(f - 1, j)];
        lz[g.at(f - 1, j)] = 0;
      }
    // 5^T and 4^T: q_z = A2 * lz on the vz region, q_x = A2 * lx on the vx region
    T* qz = scratch_q.data();
    T* qx = scratch_q.data() + plane;
    std::fill(qz, qz + plane, (T)0);
    std::fill(qx, qx + plane, (T)0);
#pragma omp parallel for schedule(static)
    for (int i = f; i <= NZ - 3; ++i)
      for (int j = 1; j <= NX - 2; ++j) {
        const size_t q = g.at(i, j);
        qz[q] = k.a2[q] * lz[q];
        stepA2[q] -= lz[q] * vz_bracket(g, P1, i, j);
        lz[q] = k.omk3[q] * lz[q];
      }
#pragma omp parallel for schedule(static)
    for (int i = f; i <= NZ - 2; ++i)
      for (int j = 1; j <= NX - 3; ++j) {
        const size_t q = g.at(i, j);
        qx[q] = k.a2[q] * lx[q];
        stepA2[q] -= lx[q] * vx_bracket(g, P1, i, j);
        lx[q] = k.omk2[q] * lx[q];
      }
    // lambda_p' += -G^T q  (gather form)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NZ; ++i)
      for (int j = 0; j < NX; ++j) {
        T acc = 0;
        auto QZ = [&](int a) -> T { return (a >= 0 && a < NZ) ? qz[g.at(a, j)] : (T)0; };
        auto QX = [&](int b) -> T { return (b >= 0 && b < NX) ? qx[g.at(i, b)] : (T)0; };
        acc += c1 * (QZ(i) - QZ(i - 1)) + c2 * (QZ(i + 1) - QZ(i - 2));
        acc += c1 * (QX(j) - QX(j - 1)) + c2 * (QX(j + 1) - QX(j - 2));
        lp[g.at(i, j)] += acc;
      }
    // 3^T: p[f-1] = -p[f+1]
    if (g.fs)
      for (int j = 0; j < NX; ++j) {
        lp[g.at(f + 1, j)] -= lp[g.at(f - 1, j)];
        lp[g.at(f - 1, j)] = 0;
      }
    // 2^T: wavelet gradient
    if (gw)
      for (int k = 0; k < geo.K; ++k) {
        const T v = g.dt * lp[g.at(geo.sz[s * geo.K + k], geo.sx[s * geo.K + k])];
        if (wavelet_shared) gw[it] += v;
        else gw[(size_t)geo.wrow(s, k) * geo.nt + it] += v;
      }
    // 1^T: r = A1 * lambda_p' on the p region; imaging condition
    T* r = qz;  // reuse
    std::fill(r, r + plane, (T)0);
#pragma omp parallel for schedule(static)
    for (int i = f + 1; i <= NZ - 3; ++i)
      for (int j = 2; j <= NX - 3; ++j) {
        const size_t q = g.at(i, j);
        r[q] = k.a1[q] * lp[q];
        stepA1[q] -= lp[q] * p_bracket(g, VX0, VZ0, i, j);
        lp[q] = k.omk1[q] * lp[q];
      }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NZ; ++i)
      for (int j = 0; j < NX; ++j) {
        auto RR = [&](int a, int b) -> T {
          return (a >= 0 && a < NZ && b >= 0 && b < NX) ? r[g.at(a, b)] : (T)0;
        };
        const size_t q = g.at(i, j);
        if (j <= NX - 2) lx[q] += c1 * (RR(i, j + 1) - RR(i, j)) + c2 * (RR(i, j + 2) - RR(i, j - 1));
        if (i <= NZ - 2) lz[q] += c1 * (RR(i + 1, j) - RR(i, j)) + c2 * (RR(i + 2, j) - RR(i - 1, j));
      }
  }
  for (size_t q = 0; q < plane; ++q) {
    gA1[q] += stepA1[q];
    gA2[q] += stepA2[q];
  }
}

// bf16 checkpoint emulation: per (field, shot) plane a power-of-two scale
// 2^(E-126) >= max|x| (E = biased exponent of max|x|), then round-to-nearest-
// even of the FP32 bit pattern to 8 significant bits.  Defines what the GPU
// library's bf16 checkpoints hold.
static float bf16_round(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

template <typename T>
static void bf16_checkpoint(std::vector<T>& fld, size_t plane, int S) {
  for (int s = 0; s < S; ++s) {
    T* p = fld.data() + (size_t)s * plane;
    float m = 0.f;
    for (size_t i = 0; i < plane; ++i) m = std::max(m, std::fabs((float)p[i]));
    unsigned mb;
    std::memcpy(&mb, &m, 4);
    const unsigned e = (mb >> 23) & 0xffu;
    float scale = 1.f;
    if (m > 0.f && e < 254u) {
      const unsigned sb = (e + 1u) << 23;
      std::memcpy(&scale, &sb, 4);
    }
    const float inv = 1.f / scale;
    for (size_t i = 0; i < plane; ++i) p[i] = (T)(bf16_round((float)p[i] * inv) * scale);
  }
}

template <typename T>
static int gradient_impl(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                         const int64_t* src_z, const int64_t* src_x, const float* src_v,
                         int wavelet_shared, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                         const float* vp, const float* rho, const float* damp, int chunk_len,
                         const T* g_p, const T* g_vx, const T* g_vz, T* grad_vp, T* grad_rho,
                         T* grad_w, int ckpt_bf16, int K = 1, int per_point = 0) {
  Grid<T> g; Geometry<T> geo;
  int rc = setup<T>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, R, rcv_z, rcv_x, g, geo,
                    K, per_point);
  if (rc) return rc;
  Coefs<T> k; build_coefs(g, vp, rho, damp, k);
  const size_t plane = g.cells();
  const size_t n = (size_t)S * plane;
  if (chunk_len <= 0 || chunk_len > nt) chunk_len = nt;
  std::vector<int> bounds = chunk_starts(nt, chunk_len);
  const int nseg = (int)bounds.size() - 1;
  // forward sweep keeping the state at the start of every segment
  std::vector<State<T>> ckpt(nseg);
  State<T> st; st.init(n);
  for (int c = 0; c < nseg; ++c) {
    ckpt[c] = st;
    if (ckpt_bf16) {
      bf16_checkpoint(ckpt[c].p, plane, S);
      bf16_checkpoint(ckpt[c].vx, plane, S);
      bf16_checkpoint(ckpt[c].vz, plane, S);
    }
    for (int it = bounds[c]; it < bounds[c + 1]; ++it) step_forward(g, k, geo, it, st);
  }
  // adjoint sweep
  State<T> lam; lam.init(n);
  std::vector<T> gA1(plane, (T)0), gA2(plane, (T)0), scratch(2 * plane);
  if (grad_w)
    std::fill(grad_w, grad_w + (wavelet_shared ? (size_t)nt : (size_t)S * (per_point ? K : 1) * nt),
              (T)0);
  for (int c = nseg - 1; c >= 0; --c) {
    const int t0 = bounds[c], t1 = bounds[c + 1];
    std::vector<State<T>> seg(t1 - t0 + 1);
    seg[0] = ckpt[c];
    for (int it = t0; it < t1; ++it) {
      seg[it - t0 + 1] = seg[it - t0];
      step_forward(g, k, geo, it, seg[it - t0 + 1]);
    }
    for (int it = t1 - 1; it >= t0; --it) {
      // records of step `it` are read from state it+1: inject their adjoint
      for (int s = 0; s < S; ++s)
        for (int r = 0; r < R; ++r) {
          const size_t q = s * plane + g.at(geo.rz[r], geo.rx[r]);
          const size_t o = ((size_t)s * nt + it) * R + r;
          if (g_p) lam.p[q] += g_p[o];
          if (g_vx) lam.vx[q] += g_vx[o];
          if (g_vz) lam.vz[q] += g_vz[o];
        }
      step_adjoint(g, k, geo, it, seg[it - t0], seg[it - t0 + 1], lam, gA1, gA2, grad_w,
                   wavelet_shared, scratch);
    }
  }
  // chain rule: alpha1 = den*c*c*dt/dz, alpha2 = dt/(den*dz), then fold pad_model
  if (grad_vp) std::fill(grad_vp, grad_vp + (size_t)nz * nx, (T)0);
  if (grad_rho) std::fill(grad_rho, grad_rho + (size_t)nz * nx, (T)0);
  for (int i = 0; i < g.NZ; ++i) {
    const int a = clampi(i - nabc, 0, nz - 1);
    for (int j = 0; j < g.NX; ++j) {
      const int b = clampi(j - nabc, 0, nx - 1);
      const size_t q = g.at(i, j);
      const T c = k.c[q], den = k.den[q];
      if (grad_vp) grad_vp[(size_t)a * nx + b] += gA1[q] * ((T)2 * den * c * g.dt / g.dz);
      if (grad_rho)
        grad_rho[(size_t)a * nx + b] +=
            gA1[q] * (c * c * g.dt / g.dz) - gA2[q] * (g.dt / (den * den * g.dz));
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int oracle_forward_f32(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                       const int64_t* src_z, const int64_t* src_x, const float* src_v, int R,
                       const int64_t* rcv_z, const int64_t* rcv_x, const float* vp,
                       const float* rho, const float* damp, int chunk_len, float* rp, float* rvx,
                       float* rvz, float* ep, float* evx, float* evz, float* snapshots,
                       float* final_state) {
  return forward_impl<float>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, R, rcv_z, rcv_x,
                             vp, rho, damp, chunk_len, rp, rvx, rvz, ep, evx, evz, snapshots,
                             final_state);
}

int oracle_forward_f64(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                       const int64_t* src_z, const int64_t* src_x, const float* src_v, int R,
                       const int64_t* rcv_z, const int64_t* rcv_x, const float* vp,
                       const float* rho, const float* damp, int chunk_len, double* rp,
                       double* rvx, double* rvz, double* ep, double* evx, double* evz,
                       double* snapshots, double* final_state) {
  return forward_impl<double>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, R, rcv_z,
                              rcv_x, vp, rho, damp, chunk_len, rp, rvx, rvz, ep, evx, evz,
                              snapshots, final_state);
}

int oracle_gradient_f32(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                        const int64_t* src_z, const int64_t* src_x, const float* src_v,
                        int wavelet_shared, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                        const float* vp, const float* rho, const float* damp, int chunk_len,
                        const float* g_p, const float* g_vx, const float* g_vz, float* grad_vp,
                        float* grad_rho, float* grad_w, int ckpt_bf16) {
  return gradient_impl<float>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v,
                              wavelet_shared, R, rcv_z, rcv_x, vp, rho, damp, chunk_len, g_p,
                              g_vx, g_vz, grad_vp, grad_rho, grad_w, ckpt_bf16);
}

int oracle_gradient_f64(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                        const int64_t* src_z, const int64_t* src_x, const float* src_v,
                        int wavelet_shared, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                        const float* vp, const float* rho, const float* damp, int chunk_len,
                        const double* g_p, const double* g_vx, const double* g_vz,
                        double* grad_vp, double* grad_rho, double* grad_w, int ckpt_bf16) {
  return gradient_impl<double>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v,
                               wavelet_shared, R, rcv_z, rcv_x, vp, rho, damp, chunk_len, g_p,
                               g_vx, g_vz, grad_vp, grad_rho, grad_w, ckpt_bf16);
}

// several source points per shot: src_z / src_x hold S*K points (point k of
// shot s at s*K+k), src_v a row per shot (per_point 0) or per point (1)
int oracle_forward_mp_f32(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                          int K, int per_point, const int64_t* src_z, const int64_t* src_x,
                          const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                          const float* vp, const float* rho, const float* damp, int chunk_len,
                          float* rp, float* rvx, float* rvz, float* ep, float* evx, float* evz) {
  return forward_impl<float>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, R, rcv_z, rcv_x,
                             vp, rho, damp, chunk_len, rp, rvx, rvz, ep, evx, evz, nullptr,
                             nullptr, K, per_point);
}

int oracle_forward_mp_f64(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                          int K, int per_point, const int64_t* src_z, const int64_t* src_x,
                          const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                          const float* vp, const float* rho, const float* damp, int chunk_len,
                          double* rp, double* rvx, double* rvz, double* ep, double* evx,
                          double* evz) {
  return forward_impl<double>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, R, rcv_z,
                              rcv_x, vp, rho, damp, chunk_len, rp, rvx, rvz, ep, evx, evz,
                              nullptr, nullptr, K, per_point);
}

int oracle_gradient_mp_f32(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                           int K, int per_point, const int64_t* src_z, const int64_t* src_x,
                           const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                           const float* vp, const float* rho, const float* damp, int chunk_len,
                           const float* g_p, const float* g_vx, const float* g_vz, float* grad_vp,
                           float* grad_rho, float* grad_w) {
  return gradient_impl<float>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, 0, R, rcv_z,
                              rcv_x, vp, rho, damp, chunk_len, g_p, g_vx, g_vz, grad_vp, grad_rho,
                              grad_w, 0, K, per_point);
}

int oracle_gradient_mp_f64(int nz, int nx, int nabc, int fs, double dt, double dz, int nt, int S,
                           int K, int per_point, const int64_t* src_z, const int64_t* src_x,
                           const float* src_v, int R, const int64_t* rcv_z, const int64_t* rcv_x,
                           const float* vp, const float* rho, const float* damp, int chunk_len,
                           const double* g_p, const double* g_vx, const double* g_vz,
                           double* grad_vp, double* grad_rho, double* grad_w) {
  return gradient_impl<double>(nz, nx, nabc, fs, dt, dz, nt, S, src_z, src_x, src_v, 0, R, rcv_z,
                               rcv_x, vp, rho, damp, chunk_len, g_p, g_vx, g_vz, grad_vp,
                               grad_rho, grad_w, 0, K, per_point);
}

}  // extern "C"
