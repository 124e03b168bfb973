// gfx950 kernels of the td1c orbital Hamiltonian product.
//
// Layouts (complex values are double2 = (re, im), or float2 for the theta-grid
// and pair arrays of the mixed-precision path; radial index fastest):
//   orbitals, l-space   : [orbital][l][r]      (td1c's wfn layout)
//   orbitals, theta-grid: [orbital][theta][r]
//   pair densities / potentials: [pair][L][r] and [pair][theta][r]
// Kernels are launched from api.hip; every launch is sized on the host from
// the same constants the kernels use for their bounds checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "devmath.hpp"
#include "kernels.hpp"

namespace td1c {


// ---------------------------------------------------------------------------
// Legendre transforms along the middle axis (bas_sph2ang1/ang2sph1/
// ang2sph2/sph2ang2, src/bas/*.F90): out[b][c][r] = sum_k in_b[k][r] M_b[k][c]
// with a real K x C matrix per batch entry (selected by the m value), stored
// with its columns padded to Cp (a multiple of 32, zeros beyond C).
// MODE 0: input read directly (optionally copied to `copy`); MODE 1: input
// formed on the fly as the pair density conj(psi_j) psi_i on the theta grid
// (hprod_mkrho2p fused in).
// Block: 64 radial lanes; its 4 wavefronts form 4/KS column groups of CPW
// columns, and KS wavefronts per group split the contraction index (the
// first adds the parts through LDS).  The matrix row segment is
// wave-uniform and is read with scalar loads.
// ---------------------------------------------------------------------------
template <int MODE, int CPW, int KS, class TI, class TO>
__global__ void __launch_bounds__(256) k_transform(const TI* __restrict__ in, TO* __restrict__ out,
                                                   const double* __restrict__ mats, const float* __restrict__ matsf,
                                                   const TransformBatch* __restrict__ batch, int nr, int K, int C,
                                                   int Cp, double2* __restrict__ copy) {
  // complex64 arithmetic only when both sides are complex64
  using A2 = typename std::conditional<std::is_same<TI, float2>::value && std::is_same<TO, float2>::value, float2,
                                       double2>::type;
  using AR = typename CReal<A2>::T;
  constexpr int NG = 4 / KS;  // column groups per block
  // logical tile order: column block fastest, so the column blocks that
  // re-read one input tile run on one XCD
  const int ny = gridDim.y, nx = gridDim.x;
  const int t = xcd_remap(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z), nx * ny * gridDim.z);
  const int by = t % ny, bx = (t / ny) % nx, bz = t / (ny * nx);
  const TransformBatch bt = batch[bz];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int grp = wave % NG, ks = wave / NG;
  const int c0 = (by * NG + grp) * CPW;
  const int r = bx * 64 + lane;
  // KS wavefronts split the contraction index; the first of them adds the parts
  const int kc = (K + KS - 1) / KS;
  const int k0 = __builtin_amdgcn_readfirstlane(ks * kc);
  const int k1 = min(K, k0 + kc);
  const bool col_ok = c0 < C;
  const AR* __restrict__ M;
  if constexpr (std::is_same<AR, float>::value) M = matsf + (size_t)bt.mat * K * Cp + (col_ok ? c0 : 0);
  else M = mats + (size_t)bt.mat * K * Cp + (col_ok ? c0 : 0);
  A2 acc[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) acc[c] = cz<A2>();
  const bool row_ok = r < bt.rows;
  if (row_ok && col_ok) {
    const TI* __restrict__ x0 = in + (size_t)bt.in0 + r;
    const TI* __restrict__ x1 = in + (size_t)bt.in1 + r;
#pragma unroll 8
    for (int k = k0; k < k1; ++k) {
      A2 x;
      if (MODE == 0) {
        const TI xv = x0[(size_t)k * nr];
        x = cto<A2>(xv);
        if constexpr (std::is_same<TI, double2>::value)
          if (copy != nullptr && c0 == 0) copy[(size_t)bt.in0 + (size_t)k * nr + r] = xv;
      } else {
        x = cmulc(cto<A2>(x0[(size_t)k * nr]), cto<A2>(x1[(size_t)k * nr]));
      }
      const AR* __restrict__ mrow = M + (size_t)k * Cp;
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        const AR m = mrow[c];
        acc[c].x = fma(x.x, m, acc[c].x);
        acc[c].y = fma(x.y, m, acc[c].y);
      }
    }
  }
  if (KS > 1) {
    __shared__ A2 red[KS > 1 ? KS - 1 : 1][NG][CPW][64];
    if (ks > 0) {
#pragma unroll
      for (int c = 0; c < CPW; ++c) red[ks - 1][grp][c][lane] = acc[c];
    }
    __syncthreads();
    if (ks == 0) {
#pragma unroll
      for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int j = 0; j < KS - 1; ++j) acc[c] = cadd(acc[c], red[j][grp][c][lane]);
    }
  }
  if (ks == 0 && row_ok && col_ok) {
#pragma unroll
    for (int c = 0; c < CPW; ++c)
      if (c0 + c < C) out[(size_t)bt.out + (size_t)(c0 + c) * nr + r] = cto<TO>(acc[c]);
  }
}

// ---------------------------------------------------------------------------
// Velocity-gauge frozen core on the theta grid (hprod_fcorbg,
// src/hprod/hprod_orbin.F90:68-105): orbg_fc = orbg0_fc * exp(-i A r cos(theta)),
// one thread per (r, theta) point.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void fc_phase_point(const FcPhaseJob& j, int r, int th) {
  if (r >= j.nr || th >= j.nlat) return;
  double s, c;
  sincos(-j.lf[8] * j.xr[r] * j.cost[th], &s, &c);
  T* orbg = static_cast<T*>(j.orbg);
  for (int f = 0; f < j.nfc; ++f) {
    const size_t q = ((size_t)f * j.nlat + th) * j.nr + r;
    orbg[q] = cto<T>(r < j.nradfc ? cmul(j.orbg0[q], make_double2(c, s)) : make_double2(0.0, 0.0));
  }
}

template <class T>
__global__ void k_fc_phase(const FcPhaseJob j) {
  fc_phase_point<T>(j, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// The same transform on the matrix cores.  Per batch entry it is the real
// GEMM Out^T = M^T X, with X[k][r] the complex input: its real and imaginary
// parts are two MFMA B operands sharing the A operand (the real matrix).
// Block: one tile of 16 radial rows and up to 4 column tiles of 16, one
// wavefront per column tile, each running the whole contraction in steps of
// 4 (MODE 1: over the pair products staged in LDS).  Lane layouts of the 16x16x4 MFMAs on
// gfx950: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k; register v of
// the result in lane l holds column l % 16 and row 4 (l / 16) + v (f32) or
// l / 16 + 4 v (f64).
// ---------------------------------------------------------------------------
template <class AR> struct Mfma16;
template <> struct Mfma16<float> {
  typedef float V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 run(float a, float b, V4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Mfma16<double> {
  typedef double V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V4 run(double a, double b, V4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <int MODE, class TI, class TO>
__device__ __forceinline__ void transform_mfma_body(const TI* __restrict__ in, TO* __restrict__ out, const double* __restrict__ mats, const float* __restrict__ matsf, const TransformBatch* __restrict__ batch, int nr, int K, int C, int Cp, double2* __restrict__ copy, int nx, int ny, int ntile, const FcPhaseJob& fc, const Blk bk) {
  using A2 = typename std::conditional<std::is_same<TI, float2>::value && std::is_same<TO, float2>::value, float2,
                                       double2>::type;
  using AR = typename CReal<A2>::T;
  using MF = Mfma16<AR>;
  using V4 = typename MF::V4;
  constexpr int KC = 64;  // contraction values staged in LDS per pass
  // blocks past the ntile transform tiles run the frozen-core phase job
  if ((int)bk.x >= ntile) {
    const int q = (int)bk.x - ntile, rb = (fc.nr + blockDim.x - 1) / blockDim.x;
    fc_phase_point<TO>(fc, (q % rb) * blockDim.x + threadIdx.x, q / rb);
    return;
  }
  __shared__ A2 sX[MODE == 1 ? KC * 16 : 1];
  const int t = xcd_remap(bk.x, ntile);
  const int by = t % ny, bx = (t / ny) % nx, bz = t / (ny * nx);
  const TransformBatch bt = batch[bz];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = (by * (int)(blockDim.x >> 6) + wave) * 16;  // first column of this wavefront's tile
  const bool active = c0 < C;  // wave-uniform
  const int r0 = bx * 16;
  const int kq = lane >> 4;
  const int ca = c0 + (lane & 15);  // matrix column this lane feeds to A
  const bool ca_ok = ca < Cp;       // columns C..Cp-1 of the padded matrix are zero
  const AR* __restrict__ M;
  if constexpr (std::is_same<AR, float>::value) M = matsf + (size_t)bt.mat * K * Cp;
  else M = mats + (size_t)bt.mat * K * Cp;
  V4 dre = {0, 0, 0, 0}, dim = {0, 0, 0, 0};
  if constexpr (MODE == 1) {
    // the pair products, formed once per block into LDS for all its
    // column-tile wavefronts
    for (int k0 = 0; k0 < K; k0 += KC) {
      // stage: element q of the tile is (k = k0 + q / 16, row r0 + q % 16),
      // four elements per thread per round with the loads issued together
      for (int q0 = 0; q0 < KC * 16; q0 += 4 * (int)blockDim.x) {
        TI x0[4], x1[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int q = q0 + u * (int)blockDim.x + (int)threadIdx.x;
          const int k = k0 + (q >> 4), r = r0 + (q & 15);
          ok[u] = q < KC * 16 && k < K && r < bt.rows;
          const size_t g = (size_t)k * nr + r;
          x0[u] = ok[u] ? in[(size_t)bt.in0 + g] : TI{};
          x1[u] = ok[u] ? in[(size_t)bt.in1 + g] : TI{};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int q = q0 + u * (int)blockDim.x + (int)threadIdx.x;
          if (q < KC * 16) sX[q] = ok[u] ? cmulc(cto<A2>(x0[u]), cto<A2>(x1[u])) : cz<A2>();
        }
      }
      __syncthreads();
      if (active) {
        const int nks = (min(KC, K - k0) + 3) >> 2;
#pragma unroll 4
        for (int st = 0; st < nks; ++st) {
          const int kl = 4 * st + kq;
          const int k = k0 + kl;
          const AR a = (k < K && ca_ok) ? M[(size_t)k * Cp + ca] : AR(0);
          const A2 x = sX[kl * 16 + (lane & 15)];
          dre = MF::run(a, x.x, dre);
          dim = MF::run(a, x.y, dim);
        }
      }
      __syncthreads();
    }
  } else if (active) {
    // plain input: each wavefront reads its B operand directly (the other
    // column tiles of the block hit the same lines in L1)
    constexpr int U = 4;  // contraction steps per load batch
    const int r = r0 + (lane & 15);
    const bool row_ok = r < bt.rows;
    const TI* __restrict__ x0 = in + (size_t)bt.in0 + r;
    const int nks = (K + 3) >> 2;
    for (int s0 = 0; s0 < nks; s0 += U) {
      AR a[U];
      A2 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = 4 * (s0 + u) + kq;
        const bool kok = k < K;
        a[u] = (kok && ca_ok) ? M[(size_t)k * Cp + ca] : AR(0);
        x[u] = cz<A2>();
        if (kok && row_ok) {
          const TI xv = x0[(size_t)k * nr];
          x[u] = cto<A2>(xv);
          if constexpr (std::is_same<TI, double2>::value)
            if (copy != nullptr && c0 == 0) copy[(size_t)bt.in0 + (size_t)k * nr + r] = xv;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        dre = MF::run(a[u], x[u].x, dre);
        dim = MF::run(a[u], x[u].y, dim);
      }
    }
  }
  const int r = r0 + (lane & 15);
  if (active && r < bt.rows) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = c0 + MF::row(lane, v);
      A2 o;
      o.x = dre[v];
      o.y = dim[v];
      if (c < C) out[(size_t)bt.out + (size_t)c * nr + r] = cto<TO>(o);
    }
  }
}

template <int MODE, class TI, class TO>
__global__ void __launch_bounds__(256) k_transform_mfma(const TI* __restrict__ in, TO* __restrict__ out,
                                                        const double* __restrict__ mats,
                                                        const float* __restrict__ matsf,
                                                        const TransformBatch* __restrict__ batch, int nr, int K, int C,
                                                        int Cp, double2* __restrict__ copy, int nx, int ny, int ntile,
                                                        const FcPhaseJob fc) {
  transform_mfma_body<MODE, TI, TO>(in, out, mats, matsf, batch, nr, K, C, Cp, copy, nx, ny, ntile, fc, hw_blk());
}

// The matrix-core transform with register blocking over columns: a
// wavefront computes 16 radial rows x 16 CT columns, so that each input
// value it loads (the MFMA B operand) feeds 2 CT MFMAs.  Block: four
// wavefronts on consecutive row tiles sharing the same columns (their A
// operands hit the same L1 lines).  Each output element is accumulated in
// the same order as in k_transform_mfma, so both kernels give the same bits.
template <int CT, int U, class TI, class TO>
__device__ __forceinline__ void transform_mfma_w_body(const TI* __restrict__ in, TO* __restrict__ out, const double* __restrict__ mats, const float* __restrict__ matsf, const TransformBatch* __restrict__ batch, int nr, int K, int C, int Cp, double2* __restrict__ copy, int nx, int ny, int ntile, const FcPhaseJob& fc, const Blk bk) {
  using A2 = typename std::conditional<std::is_same<TI, float2>::value && std::is_same<TO, float2>::value, float2,
                                       double2>::type;
  using AR = typename CReal<A2>::T;
  using MF = Mfma16<AR>;
  using V4 = typename MF::V4;
  if ((int)bk.x >= ntile) {
    const int q = (int)bk.x - ntile, rb = (fc.nr + blockDim.x - 1) / blockDim.x;
    fc_phase_point<TO>(fc, (q % rb) * blockDim.x + threadIdx.x, q / rb);
    return;
  }
  const int t = xcd_remap(bk.x, ntile);
  const int by = t % ny, bx = (t / ny) % nx, bz = t / (ny * nx);
  const TransformBatch bt = batch[bz];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c0 = by * 16 * CT;                   // first column of this wavefront's tiles
  const int nct = min(CT, (C - c0 + 15) >> 4);   // its column tiles (wave-uniform)
  const int r0 = (bx * 4 + wave) * 16;
  const int kq = lane >> 4;
  const int r = r0 + (lane & 15);
  const bool row_ok = r < bt.rows;
  const AR* __restrict__ M;
  if constexpr (std::is_same<AR, float>::value) M = matsf + (size_t)bt.mat * K * Cp;
  else M = mats + (size_t)bt.mat * K * Cp;
  const AR* __restrict__ Mc = M + c0 + (lane & 15);
  V4 dre[CT], dim[CT];
#pragma unroll
  for (int q = 0; q < CT; ++q) {
    dre[q] = V4{0, 0, 0, 0};
    dim[q] = V4{0, 0, 0, 0};
  }
  const int nks = (K + 3) >> 2;  // contraction steps, U per load batch
  for (int s0 = 0; s0 < nks; s0 += U) {
    AR a[U][CT];
    A2 x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 4 * (s0 + u) + kq;
      const bool kok = k < K;
#pragma unroll
      for (int q = 0; q < CT; ++q) a[u][q] = (kok && q < nct) ? Mc[(size_t)k * Cp + 16 * q] : AR(0);
      x[u] = cz<A2>();
      if (kok && row_ok) {
        const size_t g = (size_t)k * nr + r;
        const TI xv = in[(size_t)bt.in0 + g];
        x[u] = cto<A2>(xv);
        if constexpr (std::is_same<TI, double2>::value)
          if (copy != nullptr && by == 0) copy[(size_t)bt.in0 + g] = xv;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < CT; ++q) {
        dre[q] = MF::run(a[u][q], x[u].x, dre[q]);
        dim[q] = MF::run(a[u][q], x[u].y, dim[q]);
      }
  }
  if (row_ok) {
#pragma unroll
    for (int q = 0; q < CT; ++q) {
      if (q >= nct) break;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int c = c0 + 16 * q + MF::row(lane, v);
        A2 o;
        o.x = dre[q][v];
        o.y = dim[q][v];
        if (c < C) out[(size_t)bt.out + (size_t)c * nr + r] = cto<TO>(o);
      }
    }
  }
}

template <int CT, int U, class TI, class TO>
__global__ void __launch_bounds__(256) k_transform_mfma_w(const TI* __restrict__ in, TO* __restrict__ out,
                                                          const double* __restrict__ mats,
                                                          const float* __restrict__ matsf,
                                                          const TransformBatch* __restrict__ batch, int nr, int K,
                                                          int C, int Cp, double2* __restrict__ copy, int nx, int ny,
                                                          int ntile, const FcPhaseJob fc) {
  transform_mfma_w_body<CT, U, TI, TO>(in, out, mats, matsf, batch, nr, K, C, Cp, copy, nx, ny, ntile, fc, hw_blk());
}

// ---------------------------------------------------------------------------
// One-body part h01 = (T_l + h1) psi (hprod_tprodp, src/hprod/hprod_tprod.F90:152-235;
// hprod_zprodp, hprod_zprod.F90:108-135; hprod_pzprodp, hprod_pzprod.F90:114-213).
// Thread per (r, l, orbital).  The FEDVR band is walked with the transposed
// row bands band[d][r], d = j - r + n.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_onebody(const double2* __restrict__ orb, double2* __restrict__ h01,
                                                 const OneBodyArgs a) {
  const double2 zf = make_double2(a.lf[2], 0.0), ef = make_double2(a.lf[5], 0.0);
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  const int f = a.f0 + blockIdx.z;
  if (r >= a.nr || l >= a.L1 || f >= a.f1) return;
  const int m = a.mval[f];
  const int am = m < 0 ? -m : m;
  const bool fc = f < a.nfcore;
  const int rows = fc ? a.nradfc : a.nr;
  double2* out = h01 + ((size_t)f * a.L1 + l) * a.nr + r;
  if (r >= rows || l < am) {
    *out = make_double2(0.0, 0.0);
    return;
  }
  const int n = a.ndvr;
  const double2* psi_l = orb + ((size_t)f * a.L1 + l) * a.nr;
  const double* band = l == 0 ? a.band_k : a.band_k0;
  const int jlo = max(0, r - n), jhi = min(a.nr - 1, r + n);
  const double xr = a.xr[r];
  const double* al = a.alph + (size_t)(m + a.mmax1) * a.L1;
  double2 h = make_double2(0.0, 0.0);
  // iprojfc = 1: frozen core gets only the velocity-gauge E.z term
  // (hprod_htot.F90:100-124)
  if (fc && !a.fc_full) {
    if (a.igauge == 1) {
      if (l + 1 < a.L1) cfma(h, psi_l[a.nr + r], cscale(ef, xr * al[l]));
      if (l - 1 >= am) cfma(h, psi_l[-(ptrdiff_t)a.nr + r], cscale(ef, xr * al[l - 1]));
    }
    *out = h;
    return;
  }
  // kinetic + nuclear + centrifugal (+ frozen-core Hartree) band product
  for (int j = jlo; j <= jhi; ++j) {
    int d = j - r + n;
    double2 t;
    if (d == n) {
      t = a.tdiag[(size_t)l * a.nr + r];
    } else {
      t = make_double2(band[(size_t)d * a.nr + r], 0.0);
      if (t.x == 0.0) continue;
    }
    cfma(h, psi_l[j], t);
  }
  // dipole coupling l +- 1 in the chosen gauge
  if (a.igauge == 0) {
    if (l + 1 < a.L1) {
      double2 tmp = cscale(zf, xr * al[l]);
      cfma(h, psi_l[a.nr + r], tmp);
    }
    if (l - 1 >= am) {
      double2 tmp = cscale(zf, xr * al[l - 1]);
      cfma(h, psi_l[-(ptrdiff_t)a.nr + r], tmp);
    }
  } else {
    // d psi at l+1 and l-1 with the radial derivative band
    if (l + 1 < a.L1) {
      const double2* p1 = psi_l + a.nr;
      double2 dw = make_double2(0.0, 0.0);
      for (int j = jlo; j <= jhi; ++j) {
        double pm = a.band_p[(size_t)(j - r + n) * a.nr + r];
        if (pm == 0.0) continue;
        dw.x = fma(p1[j].x, pm, dw.x);
        dw.y = fma(p1[j].y, pm, dw.y);
      }
      // tmp1 = A * (-i alph), tmp2 = A * (-i alph (l+1) / r)
      double2 tmp1 = cmul(zf, make_double2(0.0, -al[l]));
      double2 tmp2 = cmul(zf, make_double2(0.0, -(al[l] * (l + 1.0)) / xr));
      cfma(h, dw, tmp1);
      cfma(h, p1[r], tmp2);
    }
    if (l - 1 >= am) {
      const double2* p0 = psi_l - a.nr;
      double2 dw = make_double2(0.0, 0.0);
      for (int j = jlo; j <= jhi; ++j) {
        double pm = a.band_p[(size_t)(j - r + n) * a.nr + r];
        if (pm == 0.0) continue;
        dw.x = fma(p0[j].x, pm, dw.x);
        dw.y = fma(p0[j].y, pm, dw.y);
      }
      double2 tmp1 = cmul(zf, make_double2(0.0, -al[l - 1]));
      double2 tmp2 = cmul(zf, make_double2(0.0, -(al[l - 1] * (double)l) / xr));
      cfma(h, dw, tmp1);
      cfma(h, p0[r], cscale(tmp2, -1.0));
    }
    // frozen core in the velocity gauge also carries E.z (hprod_htot.F90:123)
    if (fc) {
      if (l + 1 < a.L1) cfma(h, psi_l[a.nr + r], cscale(ef, xr * al[l]));
      if (l - 1 >= am) cfma(h, psi_l[-(ptrdiff_t)a.nr + r], cscale(ef, xr * al[l - 1]));
    }
  }
  *out = h;
}

// ---------------------------------------------------------------------------
// One-body part in element-local form (same operator as k_onebody): a block
// owns GE consecutive elements of one orbital and walks lc consecutive l
// values.  Lane (element, m) keeps row m of the element's kinetic and d/dr
// matrices in registers; psi(l-1), psi(l), psi(l+1) of the block's points
// slide through LDS.  An interior point needs one row product; a bridge adds
// the last row of the element on its left (the lane after the GE elements
// recomputes it for the block's first element).  With GE = 3 the block is a
// single wavefront (63 + 1 lanes) and the bridge partials move by lane
// shuffles instead of LDS.  The (GE, lc) pair is a launch shape chosen per
// problem by the autotuner; every shape gives the same bits.  N = ndvr.
// ---------------------------------------------------------------------------
// The convergence test of a clh1itr term (k_cn_check's arithmetic) in a
// block of BS threads: the partial sums added by 256 lanes in k_cn_check's
// order (lane v takes v, v + 256, ...; a butterfly in each wave of 64
// lanes; the four wave sums in order), so that every block size decides
// with the same bits.  Every thread of the block gets the result.
template <int BS>
__device__ __forceinline__ bool cn_series_done(const double* __restrict__ part, int ntot, int nfun, double thresh) {
  constexpr int V = 256;
  static_assert(V % BS == 0 && BS % 64 == 0, "block of whole waves dividing 256");
  __shared__ double sw[V / 64];
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < V / BS; ++j) {
    const int v = t + j * BS;
    double s = 0.0;
    for (int q = v; q < ntot; q += V) s += part[q];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((v & 63) == 0) sw[v >> 6] = s;
  }
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int w = 0; w < V / 64; ++w) tot += sw[w];
  return sqrt(tot / nfun) < thresh;
}

template <int N, int GE, int CN>
__device__ __forceinline__ void onebody_el_body(const double2* __restrict__ orb, double2* __restrict__ h01, const OneBodyElArgs& a, int lc, const Blk bk) {
  if (CN == 2 && a.stop[0]) return;
  if (CN == 2 && a.chk_part) {
    constexpr int BSC = ((GE * (N + 1) + 1) + 63) / 64 * 64;
    if (cn_series_done<BSC>(a.chk_part, a.chk_n, a.chk_nfun, a.chk_thresh)) {
      if (bk.x == 0 && bk.y == 0 && bk.z == 0 && threadIdx.x == 0) *a.chk_flag = 1;
      return;
    }
  }
  const double2 zf = make_double2(a.lf[2], 0.0), ef = make_double2(a.lf[5], 0.0);
  constexpr int NP = N + 1;
  constexpr int BS = ((GE * NP + 1) + 63) / 64 * 64;
  constexpr bool ONE_WAVE = BS == 64;
  constexpr int TILE = (GE + 1) * N + 1;     // points of elements e0-1 .. e0+GE-1
  __shared__ double2 sPsi[3][TILE];
  __shared__ double2 sT[ONE_WAVE ? 1 : BS], sP1[ONE_WAVE ? 1 : BS], sP0[ONE_WAVE ? 1 : BS];
  const int f = bk.z;
  const int e0 = bk.x * GE;
  const int l0 = bk.y * lc;
  const int l1 = min(a.L1, l0 + lc);
  const int tid = threadIdx.x;
  // lane -> (element, local row)
  int e, m;
  if (tid < GE * NP) {
    e = e0 + tid / NP;
    m = tid % NP;
  } else if (tid == GE * NP) {
    e = e0 - 1;
    m = N;
  } else {
    e = -1;
    m = 0;
  }
  const bool lane_ok = e >= 0 && e < a.nfe;
  const int ee = lane_ok ? a.emap[e] : 0;  // the element's matrix block
  const int mf = a.mval[CN ? a.fshift + f : f];
  const int am = mf < 0 ? -mf : mf;
  const bool fc = CN ? false : f < a.nfcore;
  const int rows = fc ? a.nradfc : a.nr;
  const int ibase = (e0 - 1) * N - 1;        // global interior index of tile slot 0
  // tile slot of this lane's element point 0
  const int eslot = (e - (e0 - 1)) * N;
  double Kr[NP], Pr[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    Kr[k] = lane_ok ? a.eK[((size_t)ee * NP + k) * NP + m] : 0.0;
    Pr[k] = lane_ok ? a.eP[((size_t)ee * NP + k) * NP + m] : 0.0;
  }
  const int i = e * N + m - 1;               // global interior index of (e, m)
  // owner: interior rows, and row 0 of elements >= 1 (the bridge)
  const bool owner = lane_ok && tid < GE * NP && m < N && i >= 0 && i < a.nr;
  const double xr = owner ? a.xr[i] : 1.0;
  const double* al = a.alph + (size_t)(mf + a.mmax1) * a.L1;
  const double2* base_f = orb + (size_t)f * a.L1 * a.nr;
  // the lane holding the left element's last row: the previous lane, or the
  // extra lane for the first element
  const int pl = tid >= NP ? tid - 1 : GE * NP;
  auto load_tile = [&](int l, int buf) {
    for (int q = tid; q < TILE; q += BS) {
      const int gi = ibase + q;
      double2 v = make_double2(0.0, 0.0);
      if (l >= 0 && l < a.L1 && gi >= 0 && gi < a.nr) v = base_f[(size_t)l * a.nr + gi];
      sPsi[buf][q] = v;
    }
  };
  // plane l + 1 is read from global memory one trip ahead, into registers,
  // so that its latency overlaps the previous trip's products
  constexpr int RP = (TILE + BS - 1) / BS;
  double2 nx[RP];
  auto fetch = [&](int l) {
#pragma unroll
    for (int j = 0; j < RP; ++j) {
      const int q = tid + j * BS, gi = ibase + q;
      nx[j] = (q < TILE && l >= 0 && l < a.L1 && gi >= 0 && gi < a.nr) ? base_f[(size_t)l * a.nr + gi]
                                                                       : make_double2(0.0, 0.0);
    }
  };
  load_tile(l0, l0 % 3);
  load_tile(l0 - 1, (l0 + 2) % 3);
  fetch(l0 + 1);
  for (int l = l0; l < l1; ++l) {
    const int b0 = l % 3, b1 = (l + 1) % 3, bm = (l + 2) % 3;
    __syncthreads();  // the previous trip's reads of buffer b1 are done
#pragma unroll
    for (int j = 0; j < RP; ++j)
      if (tid + j * BS < TILE) sPsi[b1][tid + j * BS] = nx[j];
    if (l + 1 < l1) fetch(l + 2);
    __syncthreads();
    // element-row partial products
    double2 pT = make_double2(0.0, 0.0), pP1 = make_double2(0.0, 0.0), pP0 = make_double2(0.0, 0.0);
    if (lane_ok && l >= am && CN != 2) {
      const double2* s0 = sPsi[b0] + eslot;
      if (l == 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          const double kk = a.eK0[((size_t)ee * NP + k) * NP + m];
          pT.x = fma(kk, s0[k].x, pT.x);
          pT.y = fma(kk, s0[k].y, pT.y);
        }
      } else {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          pT.x = fma(Kr[k], s0[k].x, pT.x);
          pT.y = fma(Kr[k], s0[k].y, pT.y);
        }
      }
    }
    if (lane_ok && l >= am) {
      if (a.igauge == 1 && (CN != 1 || a.isplit == 1)) {
        const double2* s1 = sPsi[b1] + eslot;
        const double2* sm = sPsi[bm] + eslot;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          pP1.x = fma(Pr[k], s1[k].x, pP1.x);
          pP1.y = fma(Pr[k], s1[k].y, pP1.y);
          pP0.x = fma(Pr[k], sm[k].x, pP0.x);
          pP0.y = fma(Pr[k], sm[k].y, pP0.y);
        }
      }
    }
    // the left neighbour's partials (bridge rows)
    double2 lT, lP1, lP0;
    if (ONE_WAVE) {
      lT = make_double2(__shfl(pT.x, pl, 64), __shfl(pT.y, pl, 64));
      lP1 = make_double2(__shfl(pP1.x, pl, 64), __shfl(pP1.y, pl, 64));
      lP0 = make_double2(__shfl(pP0.x, pl, 64), __shfl(pP0.y, pl, 64));
    } else {
      sT[tid] = pT;
      sP1[tid] = pP1;
      sP0[tid] = pP0;
      __syncthreads();
      lT = sT[pl];
      lP1 = sP1[pl];
      lP0 = sP0[pl];
    }
    if (owner) {
      double2 out = make_double2(0.0, 0.0);
      if (l >= am && i < rows) {
        double2 T = pT, P1 = pP1, P0 = pP0;
        if (m == 0) {  // bridge: add the left element's last row
          T = cadd(T, lT);
          P1 = cadd(P1, lP1);
          P0 = cadd(P0, lP0);
        }
        const int slot = eslot + m;
        const double2 p0 = sPsi[b0][slot], pu = sPsi[b1][slot], pd = sPsi[bm][slot];
        double2 h = make_double2(0.0, 0.0);
        const bool full = !fc || a.fc_full;
        if (CN) {
          // the clh1itr right-hand sides: T psi (CN 1), then V psi (CN 1 with
          // isplit = 1, and CN 2), as h1itr_prop.F90:57-70,106-114; CN 3
          // (T + V) psi, hprod_h1tot (hprod_h1tot.F90:2-53)
          if (CN == 1 || CN == 3) {
            h = T;
            cfma(h, p0, a.tzor[(size_t)l * a.nr + i]);
          }
          if (CN >= 2 || a.isplit == 1) {
            if (a.igauge == 0) {
              if (l + 1 < a.L1) cfma(h, pu, cscale(zf, xr * al[l]));
              if (l - 1 >= am) cfma(h, pd, cscale(zf, xr * al[l - 1]));
            } else {
              if (l + 1 < a.L1) {
                cfma(h, P1, cmul(zf, make_double2(0.0, -al[l])));
                cfma(h, pu, cmul(zf, make_double2(0.0, -(al[l] * (l + 1.0)) / xr)));
              }
              if (l - 1 >= am) {
                cfma(h, P0, cmul(zf, make_double2(0.0, -al[l - 1])));
                cfma(h, pd, cscale(cmul(zf, make_double2(0.0, -(al[l - 1] * (double)l) / xr)), -1.0));
              }
            }
          }
          double2 y = CN == 1 ? p0 : make_double2(0.0, 0.0);
          cfma(y, a.cv, h);
          h01[((size_t)f * a.L1 + l) * a.nr + i] = y;
          continue;
        }
        if (full) {
          h = T;
          cfma(h, p0, a.tzor[(size_t)l * a.nr + i]);
          if (a.igauge == 0) {
            if (l + 1 < a.L1) cfma(h, pu, cscale(zf, xr * al[l]));
            if (l - 1 >= am) cfma(h, pd, cscale(zf, xr * al[l - 1]));
          } else {
            if (l + 1 < a.L1) {
              cfma(h, P1, cmul(zf, make_double2(0.0, -al[l])));
              cfma(h, pu, cmul(zf, make_double2(0.0, -(al[l] * (l + 1.0)) / xr)));
            }
            if (l - 1 >= am) {
              cfma(h, P0, cmul(zf, make_double2(0.0, -al[l - 1])));
              cfma(h, pd, cscale(cmul(zf, make_double2(0.0, -(al[l - 1] * (double)l) / xr)), -1.0));
            }
          }
        }
        if (fc && a.igauge == 1) {  // E.z on the frozen core in the velocity gauge
          if (l + 1 < a.L1) cfma(h, pu, cscale(ef, xr * al[l]));
          if (l - 1 >= am) cfma(h, pd, cscale(ef, xr * al[l - 1]));
        }
        out = h;
      }
      h01[((size_t)f * a.L1 + l) * a.nr + i] = out;
    }
  }
}

template <int N, int GE, int CN>
__global__ void __launch_bounds__(((GE * (N + 1) + 1) + 63) / 64 * 64)
    k_onebody_el(const double2* __restrict__ orb, double2* __restrict__ h01, const OneBodyElArgs a, int lc) {
  onebody_el_body<N, GE, CN>(orb, h01, a, lc, hw_blk());
}

// ---------------------------------------------------------------------------
// Radial Poisson solve per (pair, L) (hprod_mkv2mf_poissonp,
// src/hprod/hprod_mkv2mf.F90:118-182) by element Schur complement of the
// same FEDVR operator A_L = 2K + L(L+1)/r^2, b = rho * d2invr.
// Element e has interior block A_e (inverse G_e), coupling rows C_l / C_r to
// its left / right bridge, and W_l = G_e C_l^T, W_r = G_e C_r^T.  The bridge
// unknowns solve the tridiagonal Schur system
//   S x_B = b_B - sum_e (W_r(e-side) . b_I(e)) - (W_l . b_I(e+1)),
// and the interiors follow as x_I = G_e b_I - W_l x_left - W_r x_right.
// v = (x + Q r^(L+1) sqrt(w)) * d2fac2, Q = d2fac1 * sum_r rho r^L sqrt(w)/..
// (multipole boundary term, d2rpl0/d2rpl1).
// ---------------------------------------------------------------------------
// B1: one block per (pair, L), jobs ordered L-major so consecutive blocks
// share the L tables: element dots W.b (one thread per element, the NI
// interior loads of a thread issued together), the multipole sum over all
// rows, the Schur right-hand sides, the two bidiagonal sweeps of the
// factored Schur system on thread 0 out of LDS, then the bridge rows of v.
// NI = interior points per element (0: runtime a.ni).
template <int NI, class T>
__global__ void __launch_bounds__(256) k_poisson_bridge(const T* __restrict__ rho, T* __restrict__ v,
                                                        double2* __restrict__ xb, const PoissonArgs a) {
  const int tid = threadIdx.x;
  const int job = xcd_remap(blockIdx.x, gridDim.x);  // the blocks of one L share an XCD (and its L2)
  const int L = job / a.npair, p = job % a.npair;
  const int nb = a.nfe - 1;
  const int ni = NI > 0 ? NI : a.ni;
  __shared__ double2 sS[kMaxBridge + 1];
  __shared__ double2 sCr[kMaxBridge + 1];
  __shared__ double2 sY[kMaxBridge];
  __shared__ double2 sRed[4];
  __shared__ double sPA[kMaxPcrSteps * kMaxBridge], sPG[kMaxPcrSteps * kMaxBridge];
  const int pm = a.pair_m[p];
  const size_t row = ((size_t)p * a.L2 + L) * a.nr;
  T* vp = v + row;
  double2* xbp = xb + ((size_t)p * a.L2 + L) * nb;
  if (L < (pm < 0 ? -pm : pm)) {  // uniform over the block
    for (int k = tid; k < nb; k += 256) {
      vp[(k + 1) * a.n - 1] = cz<T>();
      xbp[k] = make_double2(0.0, 0.0);
    }
    if (tid == 0) a.Q[(size_t)p * a.L2 + L] = make_double2(0.0, 0.0);
    return;
  }
  const T* rp = rho + row;
  const double* invr = a.invr + (size_t)L * a.nr;
  const double* Wl = a.Wl + (size_t)L * a.nr;
  const double* Wr = a.Wr + (size_t)L * a.nr;
  const double* rpl0 = a.rpl0 + (size_t)L * a.nr;
  // per-bridge inputs of thread k = tid (nfe - 1 <= 256), loaded up front
  double2 rb = make_double2(0.0, 0.0);
  double binv = 0.0, r1 = 0.0, f2 = 0.0;
  if (tid < nb) {
    const int gi = (tid + 1) * a.n - 1;
    rb = cscale(cto<double2>(rp[gi]), invr[gi]);
    binv = a.pcrBinv[(size_t)L * nb + tid];
    r1 = a.rpl1[(size_t)L * a.nr + gi];
    f2 = a.fac2[(size_t)L * a.nr + gi];
  }
  // the cyclic-reduction coefficients of this L, staged in LDS early so
  // their loads overlap the element dots
  {
    const int K = a.pcr_steps;
    const double* pa = a.pcrA + (size_t)L * K * nb;
    const double* pg = a.pcrG + (size_t)L * K * nb;
    for (int q = tid; q < K * nb; q += 256) {
      sPA[q] = pa[q];
      sPG[q] = pg[q];
    }
  }
  // per element: cl = W_l . b_I, cr = W_r . b_I with b = rho * invr
  for (int e = tid; e < a.nfe; e += 256) {
    const int base = e * a.n;
    double2 cl = make_double2(0.0, 0.0), cr = make_double2(0.0, 0.0);
    if (NI > 0) {
      double2 x[NI > 0 ? NI : 1];
      double iv[NI > 0 ? NI : 1], wl[NI > 0 ? NI : 1], wr[NI > 0 ? NI : 1];
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        x[i] = cto<double2>(rp[base + i]);
        iv[i] = invr[base + i];
        wl[i] = Wl[base + i];
        wr[i] = Wr[base + i];
      }
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const double2 b = cscale(x[i], iv[i]);
        cl.x = fma(wl[i], b.x, cl.x);
        cl.y = fma(wl[i], b.y, cl.y);
        cr.x = fma(wr[i], b.x, cr.x);
        cr.y = fma(wr[i], b.y, cr.y);
      }
    } else {
      for (int i = 0; i < ni; ++i) {
        const double2 b = cscale(cto<double2>(rp[base + i]), invr[base + i]);
        cl.x = fma(Wl[base + i], b.x, cl.x);
        cl.y = fma(Wl[base + i], b.y, cl.y);
        cr.x = fma(Wr[base + i], b.x, cr.x);
        cr.y = fma(Wr[base + i], b.y, cr.y);
      }
    }
    // bridge k = e - 1 (left of e) takes -cl; bridge k = e (right of e) takes -cr
    if (e >= 1) sS[e - 1] = cl;
    sCr[e] = cr;
  }
  // multipole moment over every radial row
  double2 m = make_double2(0.0, 0.0);
  for (int r = tid; r < a.nr; r += 256) {
    const double2 x = cto<double2>(rp[r]);
    m.x = fma(x.x, rpl0[r], m.x);
    m.y = fma(x.y, rpl0[r], m.y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    m.x += __shfl_xor(m.x, o, 64);
    m.y += __shfl_xor(m.y, o, 64);
  }
  if ((tid & 63) == 0) sRed[tid >> 6] = m;
  __syncthreads();
  const double2 Q = cscale(cadd(cadd(sRed[0], sRed[1]), cadd(sRed[2], sRed[3])), a.fac1[L]);
  if (tid < nb) {
    double2 sk = rb;
    sk = csub(sk, sS[tid]);
    sk = csub(sk, sCr[tid]);
    sY[tid] = sk;
  }
  __syncthreads();
  // parallel cyclic reduction of the Schur system with the per-L
  // coefficients: pcr_steps stencil sweeps over the right-hand side
  double2* cur = sY;
  double2* nxt = sS;
  const int K = a.pcr_steps;
  const double* pa = sPA;
  const double* pg = sPG;
  for (int k = 0, h = 1; k < K; ++k, h <<= 1) {
    for (int i = tid; i < nb; i += 256) {
      double2 d = cur[i];
      if (i - h >= 0) {
        const double al = pa[(size_t)k * nb + i];
        d = make_double2(fma(al, cur[i - h].x, d.x), fma(al, cur[i - h].y, d.y));
      }
      if (i + h < nb) {
        const double ga = pg[(size_t)k * nb + i];
        d = make_double2(fma(ga, cur[i + h].x, d.x), fma(ga, cur[i + h].y, d.y));
      }
      nxt[i] = d;
    }
    __syncthreads();
    double2* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (tid == 0) a.Q[(size_t)p * a.L2 + L] = Q;
  if (tid < nb) {
    const int gi = (tid + 1) * a.n - 1;
    const double2 x = cscale(cur[tid], binv);
    xbp[tid] = x;
    vp[gi] = cto<T>(make_double2((x.x + Q.x * r1) * f2, (x.y + Q.y * r1) * f2));
  }
}

// B2: block per (element, L): lanes in groups of ni, one group per pair
// (the block is sized to hold every pair when it can), lane i owning
// interior row i; G_e in LDS serves every pair.
//   x_I = G_e b_I - W_l x_left - W_r x_right,  v = (x + Q rpl1) * fac2.
template <class T>
__global__ void __launch_bounds__(1024) k_poisson_elem(const T* __restrict__ rho, T* __restrict__ v,
                                                      const double2* __restrict__ xb, const PoissonArgs a) {
  // consecutive elements of one L on one XCD: neighbours share cache lines
  const int lt = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int e = lt % gridDim.x;
  const int L = lt / gridDim.x;
  const int ni = a.ni;
  const int nt = blockDim.x;
  const int ngroup = nt / ni;
  const int nb = a.nfe - 1;
  __shared__ double sG[32 * 32];
  __shared__ double sInv[32], sWl[32], sWr[32], sR1[32], sF2[32];
  __shared__ double2 sB[1024];
  const double* G = a.G + (((size_t)L * a.gstride) + e) * ni * ni;
  const int base = e * a.n;  // interior global index of local point 1
  for (int q = threadIdx.x; q < ni * ni; q += nt) sG[q] = G[q];
  for (int q = threadIdx.x; q < ni; q += nt) {
    const size_t k = (size_t)L * a.nr + base + q;
    sInv[q] = a.invr[k];
    sWl[q] = a.Wl[k];
    sWr[q] = a.Wr[k];
    sR1[q] = a.rpl1[k];
    sF2[q] = a.fac2[k];
  }
  __syncthreads();
  const int g = threadIdx.x / ni, lane = threadIdx.x % ni;
  const bool ingroup = g < ngroup;
  for (int p0 = 0; p0 < a.npair; p0 += ngroup) {
    const int p = p0 + g;
    const bool active = ingroup && p < a.npair;
    bool valid = active;
    if (valid) {
      const int pm = a.pair_m[p];
      valid = L >= (pm < 0 ? -pm : pm);
    }
    const size_t row = ((size_t)p * a.L2 + L) * a.nr;
    if (valid) sB[threadIdx.x] = cscale(cto<double2>(rho[row + base + lane]), sInv[lane]);
    __syncthreads();
    if (active) {
      double2 out = make_double2(0.0, 0.0);
      if (valid) {
        const double2* b = sB + g * ni;
        double2 y = make_double2(0.0, 0.0);
        for (int k = 0; k < ni; ++k) {
          const double gik = sG[k * ni + lane];  // G is stored transposed
          y.x = fma(gik, b[k].x, y.x);
          y.y = fma(gik, b[k].y, y.y);
        }
        const double2* xbp = xb + ((size_t)p * a.L2 + L) * nb;
        if (e >= 1) y = csub(y, cscale(xbp[e - 1], sWl[lane]));
        if (e <= a.nfe - 2) y = csub(y, cscale(xbp[e], sWr[lane]));
        const double2 Q = a.Q[(size_t)p * a.L2 + L];
        out = make_double2((y.x + Q.x * sR1[lane]) * sF2[lane], (y.y + Q.y * sR1[lane]) * sF2[lane]);
      }
      v[row + base + lane] = cto<T>(out);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Mean-field application on the theta grid (hprod_mfprod2,
// src/hprod/hprod_mfprod2.F90: fcx1, dcj, dcx, actj, actx, act2).
// Thread per grid point (theta, r); each pair potential is read once.
// ---------------------------------------------------------------------------
template <class A2>
__device__ __forceinline__ void mfprod_body(const A2* __restrict__ orbg, const A2* __restrict__ vg, A2* __restrict__ gorbg, A2* __restrict__ v2orbg, const MfArgs& a, const Blk bk) {
  const int r = bk.x * blockDim.x + threadIdx.x;
  const int th = bk.y;
  if (r >= a.nr) return;
  const size_t ng = (size_t)a.nlat * a.nr;
  const size_t g = (size_t)th * a.nr + r;
  A2 psi[kMaxFun];
  A2 ag[kMaxFun], av[kMaxFun];
  for (int f = 0; f < a.nfun; ++f) {
    psi[f] = orbg[(size_t)f * ng + g];
    ag[f] = cz<A2>();
    av[f] = cz<A2>();
  }
  A2 d2v_dc = cz<A2>();
  A2 vact[kMaxActPairs];
  const int nfc = a.nfcore, ncore = a.ncore, na = a.nact;
  // actx needs wd1_j = sum_k psi_k (-den1(k,j)/2) for m_j = m_k
  for (int p = 0; p < a.npair; ++p) {
    const A2 v = vg[(size_t)p * ng + g];
    const int j = a.pj[p], i = a.pi[p];
    if (a.ptype[p] == 1) {  // v(dyn j, fc1 i): fcx1 uses v(i, j) = conj(v)
      if (r < a.nradfc) ag[j] = csub(ag[j], cmul(cconj(v), psi[i]));
      continue;
    }
    // dyn-dyn pair, j >= i
    const bool jdc = j < ncore, idc = i < ncore;
    if (j == i && jdc) d2v_dc = cadd(d2v_dc, cscale(v, 2.0));
    if (jdc) ag[i] = csub(ag[i], cmul(v, psi[j]));                  // dcx: jfun=j, ifun=i
    if (idc && j != i) ag[j] = csub(ag[j], cmul(cconj(v), psi[i])); // dcx: jfun=i, ifun=j
    if (j >= ncore && i >= ncore) {
      vact[(j - ncore) * na + (i - ncore)] = v;
      if (j != i) vact[(i - ncore) * na + (j - ncore)] = cconj(v);
    } else if (j >= ncore && idc) {
      // actx: v2orbg(i) += wd1(ja) v(ja, i), wd1(ja) = -1/2 sum_k den1(k,ja) psi_k
      const int ja = j - ncore;
      A2 wd = cz<A2>();
      for (int ka = 0; ka < na; ++ka)
        if (a.mval[ncore + ka] == a.mval[j]) cfma(wd, psi[ncore + ka], cscale(cto<A2>(a.den1[ka + na * ja]), -0.5));
      cfma(av[i], wd, v);
    }
  }
  for (int f = nfc; f < a.nfun; ++f) cfma(ag[f], d2v_dc, psi[f]);  // dcj
  if (ncore > nfc && na > 0) {
    // actj: d2v = sum_jk v(j,k) den1(k,j) applied to dynamical core orbitals
    A2 d = cz<A2>();
    for (int ja = 0; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka) cfma(d, vact[ja * na + ka], cto<A2>(a.den1[ka + na * ja]));
    for (int f = nfc; f < ncore; ++f) cfma(av[f], d, psi[f]);
  }
  if (na > 0 && a.nelact2 >= 2) {
    // act2: t_i = sum_j psi_j d_ij, d_ij = sum_kl den2(j,i,l,k) v(k,l);
    // v2orbg_k += sum_i rden(i,k) t_i
    A2 t[kMaxAct];
    for (int q = 0; q < na; ++q) t[q] = cz<A2>();
    for (int ia = 0; ia < na; ++ia)
      for (int ja = 0; ja <= ia; ++ja) {
        A2 d = cz<A2>();
        for (int ka = 0; ka < na; ++ka)
          for (int la = 0; la < na; ++la) {
            if (a.mval[ncore + ia] + a.mval[ncore + ka] != a.mval[ncore + ja] + a.mval[ncore + la]) continue;
            const double2 dens = a.den2[ja + na * (ia + na * (la + na * ka))];
            if (fabs(dens.x) + fabs(dens.y) == 0.0) continue;
            if (sqrt(dens.x * dens.x + dens.y * dens.y) <= 1.0e-15) continue;
            cfma(d, cto<A2>(dens), vact[ka * na + la]);
          }
        cfma(t[ia], psi[ncore + ja], d);
        if (ia > ja) cfma(t[ja], psi[ncore + ia], cconj(d));
      }
    for (int ia = 0; ia < na; ++ia)
      for (int ka = 0; ka < na; ++ka)
        if (a.mval[ncore + ia] == a.mval[ncore + ka]) cfma(av[ncore + ka], cto<A2>(a.rden[ia + na * ka]), t[ia]);
  }
  for (int f = nfc; f < a.nfun; ++f) {
    gorbg[(size_t)(f - nfc) * ng + g] = ag[f];
    v2orbg[(size_t)(f - nfc) * ng + g] = av[f];
  }
}

template <class A2>
__global__ void __launch_bounds__(256) k_mfprod(const A2* __restrict__ orbg, const A2* __restrict__ vg,
                                                A2* __restrict__ gorbg, A2* __restrict__ v2orbg,
                                                const MfArgs a) {
  mfprod_body<A2>(orbg, vg, gorbg, v2orbg, a, hw_blk());
}

// Register-resident variant of k_mfprod for a fixed orbital partition
// (NFC frozen cores, NF orbitals, NA active): every orbital and pair index is
// a compile-time constant, the pairs are visited in the order they are
// stored (dyn-dyn i <= j, then dyn-fc1).
template <int NFC, int NF, int NA, class A2>
__device__ __forceinline__ void mfprod_t_body(const A2* __restrict__ orbg, const A2* __restrict__ vg, A2* __restrict__ gorbg, A2* __restrict__ v2orbg, const MfArgs& a, const Blk bk) {
  constexpr int NCORE = NF - NA;
  constexpr int NA2 = NA > 0 ? NA * NA : 1;
  constexpr int ND = NF - NFC;
  constexpr int NDP = ND * (ND + 1) / 2;  // dynamical pairs, stored first
  const int r = bk.x * blockDim.x + threadIdx.x;
  const int th = bk.y;
  if (r >= a.nr) return;
  const size_t ng = (size_t)a.nlat * a.nr;
  const size_t g = (size_t)th * a.nr + r;
  // every load of the point is issued before any arithmetic
  A2 psi[NF], ag[NF], av[NF], vdp[NDP];
#pragma unroll
  for (int f = 0; f < NF; ++f) psi[f] = orbg[(size_t)f * ng + g];
#pragma unroll
  for (int q = 0; q < NDP; ++q) vdp[q] = vg[(size_t)q * ng + g];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    ag[f] = cz<A2>();
    av[f] = cz<A2>();
  }
  // actx weights wd1_j = -1/2 sum_k den1(k,j) psi_k (m_k = m_j)
  A2 wd[NA > 0 ? NA : 1];
#pragma unroll
  for (int ja = 0; ja < NA; ++ja) {
    wd[ja] = cz<A2>();
#pragma unroll
    for (int ka = 0; ka < NA; ++ka) cfma(wd[ja], psi[NCORE + ka], cto<A2>(a.mfc[NA * NA * NA * NA + ka + NA * ja]));
  }
  A2 d2v_dc = cz<A2>();
  A2 vact[NA2];
  int p = 0;
#pragma unroll
  for (int i = NFC; i < NF; ++i) {
#pragma unroll
    for (int j = i; j < NF; ++j) {
      const A2 v = vdp[p++];
      const bool jdc = j < NCORE, idc = i < NCORE;
      if (j == i && jdc) d2v_dc = cadd(d2v_dc, cscale(v, 2.0));
      if (jdc) ag[i] = csub(ag[i], cmul(v, psi[j]));
      if (idc && j != i) ag[j] = csub(ag[j], cmul(cconj(v), psi[i]));
      if (j >= NCORE && i >= NCORE) {
        vact[(j - NCORE) * NA + (i - NCORE)] = v;
        if (j != i) vact[(i - NCORE) * NA + (j - NCORE)] = cconj(v);
      } else if (j >= NCORE && idc) {
        cfma(av[i], wd[j - NCORE], v);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NFC; ++i) {
    if (i < a.nfcore2) continue;  // type -2 cores act through v2xfc only
#pragma unroll
    for (int j = NFC; j < NF; ++j) {
      const A2 v = vg[(size_t)(p++) * ng + g];
      if (r < a.nradfc) ag[j] = csub(ag[j], cmul(cconj(v), psi[i]));
    }
  }
#pragma unroll
  for (int f = NFC; f < NF; ++f) cfma(ag[f], d2v_dc, psi[f]);
  if (NCORE > NFC && NA > 0) {
    A2 d = cz<A2>();
#pragma unroll
    for (int ja = 0; ja < NA; ++ja)
#pragma unroll
      for (int ka = 0; ka < NA; ++ka) cfma(d, vact[ja * NA + ka], cto<A2>(a.den1[ka + NA * ja]));
#pragma unroll
    for (int f = NFC; f < NCORE; ++f) cfma(av[f], d, psi[f]);
  }
  if (NA > 0 && a.nelact2 >= 2) {
    A2 t[NA > 0 ? NA : 1];
#pragma unroll
    for (int q = 0; q < NA; ++q) t[q] = cz<A2>();
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ja = 0; ja <= ia; ++ja) {
        A2 d = cz<A2>();
#pragma unroll
        for (int ka = 0; ka < NA; ++ka)
#pragma unroll
          for (int la = 0; la < NA; ++la) cfma(d, cto<A2>(a.mfc[ja + NA * (ia + NA * (la + NA * ka))]), vact[ka * NA + la]);
        cfma(t[ia], psi[NCORE + ja], d);
        if (ia > ja) cfma(t[ja], psi[NCORE + ia], cconj(d));
      }
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
      for (int ka = 0; ka < NA; ++ka) cfma(av[NCORE + ka], cto<A2>(a.mfc[NA * NA * NA * NA + NA * NA + ia + NA * ka]), t[ia]);
  }
#pragma unroll
  for (int f = NFC; f < NF; ++f) {
    gorbg[(size_t)(f - NFC) * ng + g] = ag[f];
    v2orbg[(size_t)(f - NFC) * ng + g] = av[f];
  }
}

template <int NFC, int NF, int NA, class A2>
__global__ void __launch_bounds__(256) k_mfprod_t(const A2* __restrict__ orbg, const A2* __restrict__ vg,
                                                  A2* __restrict__ gorbg, A2* __restrict__ v2orbg,
                                                  const MfArgs a) {
  mfprod_t_body<NFC, NF, NA, A2>(orbg, vg, gorbg, v2orbg, a, hw_blk());
}

// The same products for four active orbitals and no dynamical core
// (nact = 4, nfun = nfcore + 4, ncore = nfcore), four lanes per point: lane
// q owns active orbital q and takes the act2 rows that reach it,
//   t_q = sum_{x <= q} psi_x d(q, x) + sum_{x > q} psi_x conj(d(x, q)),
//   d(i, j) = sum_{k,l} act2[j + 4 (i + 4 (l + 4 k))] vact[k][l],
// in the order of k_mfprod_t, so the results agree bit for bit; the rden
// contraction v2_q = sum_i rden[i + 4 q] t_i takes t_i from the quad by
// shuffles.  The coefficients are staged in LDS.
template <int NFC, class A2>
__global__ void __launch_bounds__(256) k_mfprod_quad(const A2* __restrict__ orbg, const A2* __restrict__ vg,
                                                     A2* __restrict__ gorbg, A2* __restrict__ v2orbg,
                                                     const MfArgs a) {
  constexpr int NA = 4, NF = NFC + NA, NCORE = NFC, NDP = NA * (NA + 1) / 2;
  __shared__ A2 s_act2[NA * NA * NA * NA], s_rden[NA * NA];
  for (int e = threadIdx.x; e < NA * NA * NA * NA; e += blockDim.x) s_act2[e] = cto<A2>(a.mfc[e]);
  if (threadIdx.x < NA * NA) s_rden[threadIdx.x] = cto<A2>(a.mfc[NA * NA * NA * NA + NA * NA + threadIdx.x]);
  const int q = threadIdx.x & 3;
  const int r = blockIdx.x * (blockDim.x >> 2) + (threadIdx.x >> 2);
  const int th = blockIdx.y;
  const bool ok = r < a.nr;
  const size_t ng = (size_t)a.nlat * a.nr;
  const size_t g = (size_t)th * a.nr + (ok ? r : 0);
  A2 psi[NF], vdp[NDP], vfc[NFC > 0 ? NFC : 1];
#pragma unroll
  for (int f = 0; f < NF; ++f) psi[f] = ok ? orbg[(size_t)f * ng + g] : cz<A2>();
#pragma unroll
  for (int p = 0; p < NDP; ++p) vdp[p] = ok ? vg[(size_t)p * ng + g] : cz<A2>();
  // the frozen-core pair (i, dynamical q) of each type -1 core i
#pragma unroll
  for (int i = 0; i < NFC; ++i)
    vfc[i] = (ok && i >= a.nfcore2) ? vg[(size_t)(NDP + (i - a.nfcore2) * NA + q) * ng + g] : cz<A2>();
  __syncthreads();
  if (!ok) return;
  A2 vact[NA * NA];
  {
    int p = 0;
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = i; j < NA; ++j) {
        const A2 v = vdp[p++];
        vact[j * NA + i] = v;
        if (j != i) vact[i * NA + j] = cconj(v);
      }
  }
  A2 ag = cz<A2>(), av = cz<A2>();
#pragma unroll
  for (int i = 0; i < NFC; ++i) {
    if (i < a.nfcore2) continue;  // type -2 cores act through v2xfc only
    if (r < a.nradfc) ag = csub(ag, cmul(cconj(vfc[i]), psi[i]));
  }
  if (a.nelact2 >= 2) {
    A2 t = cz<A2>();
#pragma unroll
    for (int x = 0; x < NA; ++x) {
      // d(q, x) for x <= q, d(x, q) for x > q
      const int ia = x <= q ? q : x, ja = x <= q ? x : q;
      A2 d = cz<A2>();
#pragma unroll
      for (int ka = 0; ka < NA; ++ka)
#pragma unroll
        for (int la = 0; la < NA; ++la) cfma(d, s_act2[ja + NA * (ia + NA * (la + NA * ka))], vact[ka * NA + la]);
      cfma(t, psi[NCORE + x], x <= q ? d : cconj(d));
    }
    const int base = threadIdx.x & 60;  // first lane of the quad within the wavefront
#pragma unroll
    for (int ia = 0; ia < NA; ++ia) {
      A2 ti;
      ti.x = __shfl(t.x, base + ia, 64);
      ti.y = __shfl(t.y, base + ia, 64);
      cfma(av, s_rden[ia + NA * q], ti);
    }
  }
  gorbg[(size_t)q * ng + g] = ag;
  v2orbg[(size_t)q * ng + g] = av;
}

// ---------------------------------------------------------------------------
// Gram products over a common index range, FP64, deterministic two-stage sum
// (the inner products of hprod_mkint1_sph, hprod_mkint2_sph, hprod_ene_dcore
// and hprod_htot_mkxmat_*):
//   S0[a][b] = sum_q opA(L_a[q]) R_b[q],  S1[a][b] = sum_q opA(L_a[q]) conj(R_b[q])
// opA = conj when conj_left.  A block owns a TA x TB tile of (a, b) and a
// slice of q; each thread keeps the tile's accumulators in registers, so each
// loaded element feeds TA or TB products.  The block result goes to
// partial[entry][which][block-in-q].
// ---------------------------------------------------------------------------
template <int TA, int TB, bool CR, class T>
__device__ __forceinline__ void gram_body(const GramArgs& a, double2* __restrict__ partial, int nqblk, const Blk bk) {
  const int ntb = (a.nR + TB - 1) / TB;
  // logical order: (a, b) tile fastest, so the tiles that read one index
  // range run on one XCD
  const int nt = bk.ny;
  const int lt = xcd_remap(bk.x + bk.nx * bk.y, bk.nx * nt);
  const int tile = lt % nt, qb = lt / nt;
  const int ta = tile / ntb, tb = tile % ntb;
  const int a0 = ta * TA, b0 = tb * TB;
  const double sa = a.conj_left ? -1.0 : 1.0;
  double2 s0[TA][TB], s1[CR ? TA : 1][CR ? TB : 1];
#pragma unroll
  for (int u = 0; u < TA; ++u)
#pragma unroll
    for (int v = 0; v < TB; ++v) {
      s0[u][v] = make_double2(0.0, 0.0);
      if (CR) s1[u][v] = make_double2(0.0, 0.0);
    }
  const T* Lp[TA];
  const T* Rp[TB];
#pragma unroll
  for (int u = 0; u < TA; ++u) Lp[u] = a0 + u < a.nL ? static_cast<const T*>(a.left[a0 + u]) : nullptr;
#pragma unroll
  for (int v = 0; v < TB; ++v) Rp[v] = b0 + v < a.nR ? static_cast<const T*>(a.right[b0 + v]) : nullptr;
  // two index values per trip, all their loads issued before the products
  const size_t stride = (size_t)nqblk * 256;
  size_t q = (size_t)qb * 256 + threadIdx.x;
  for (; q < a.len; q += 2 * stride) {
    const size_t q1 = q + stride;
    const bool has1 = q1 < a.len;
    double2 x[2][TA], y[2][TB];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const size_t qq = h == 0 ? q : (has1 ? q1 : q);
#pragma unroll
      for (int u = 0; u < TA; ++u) x[h][u] = Lp[u] ? cto<double2>(Lp[u][qq]) : make_double2(0.0, 0.0);
#pragma unroll
      for (int v = 0; v < TB; ++v) y[h][v] = Rp[v] ? cto<double2>(Rp[v][qq]) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && !has1) break;
#pragma unroll
      for (int u = 0; u < TA; ++u) {
        const double2 xu = make_double2(x[h][u].x, x[h][u].y * sa);
#pragma unroll
        for (int v = 0; v < TB; ++v) {
          cfma(s0[u][v], xu, y[h][v]);
          if (CR) cfma(s1[u][v], xu, cconj(y[h][v]));
        }
      }
    }
  }
  // block reduction.  The tile holds NV = TA*TB*(CR ? 2 : 1) complex sums =
  // ND = 2*NV doubles per lane (16, 32 or 64); a transposing butterfly over
  // ND/2 ... 1 leaves lane t with the total of double t % ND over its group
  // of ND lanes, exchanges at distance ND ... 32 join the groups, and the 4
  // waves are summed through LDS.
  constexpr int NV = TA * TB * (CR ? 2 : 1);
  constexpr int ND = 2 * NV;
  static_assert(ND == 16 || ND == 32 || ND == 64, "tile must hold 16, 32 or 64 doubles per lane");
  double v[ND];
#pragma unroll
  for (int u = 0; u < TA; ++u)
#pragma unroll
    for (int w = 0; w < TB; ++w)
#pragma unroll
      for (int c = 0; c < (CR ? 2 : 1); ++c) {
        const int e = (u * TB + w) * (CR ? 2 : 1) + c;
        const double2 t = c == 0 ? s0[u][w] : s1[CR ? u : 0][CR ? w : 0];
        v[2 * e] = t.x;
        v[2 * e + 1] = t.y;
      }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#define TD1C_GRAM_STEP(H)                                      \
  _Pragma("unroll") for (int j = 0; j < H; ++j) {              \
    const bool up = (lane & H) != 0;                           \
    const double send = up ? v[j] : v[j + H];                  \
    const double keep = up ? v[j + H] : v[j];                  \
    v[j] = keep + __shfl_xor(send, H, 64);                     \
  }
  if constexpr (ND == 64) {
    TD1C_GRAM_STEP(32)
  }
  if constexpr (ND >= 32) {
    TD1C_GRAM_STEP(16)
  }
  TD1C_GRAM_STEP(8)
  TD1C_GRAM_STEP(4)
  TD1C_GRAM_STEP(2)
  TD1C_GRAM_STEP(1)
#undef TD1C_GRAM_STEP
  if constexpr (ND == 16) v[0] += __shfl_xor(v[0], 16, 64);
  if constexpr (ND <= 32) v[0] += __shfl_xor(v[0], 32, 64);
  __shared__ double red[4][64];
  red[wave][lane] = v[0];
  __syncthreads();
  if (threadIdx.x < NV) {
    const int e = threadIdx.x;
    const int c = CR ? e % 2 : 0, uv = CR ? e / 2 : e, u = uv / TB, w = uv % TB;
    if (a0 + u < a.nL && b0 + w < a.nR) {
      double2 t;
      t.x = (red[0][2 * e] + red[1][2 * e]) + (red[2][2 * e] + red[3][2 * e]);
      t.y = (red[0][2 * e + 1] + red[1][2 * e + 1]) + (red[2][2 * e + 1] + red[3][2 * e + 1]);
      const int entry = (a0 + u) * a.nR + (b0 + w);
      partial[((size_t)entry * 2 + c) * nqblk + qb] = t;
    }
  }
}

template <int TA, int TB, bool CR, class T>
__global__ void __launch_bounds__(256) k_gram(const GramArgs a, double2* __restrict__ partial, int nqblk) {
  gram_body<TA, TB, CR, T>(a, partial, nqblk, hw_blk());
}

// one wavefront per Gram entry: strided partial sums, then a butterfly
__global__ void __launch_bounds__(256) k_gram_final(const double2* __restrict__ partial, int nblk, int nent,
                                                    double2* __restrict__ out, int with_conj) {
  const int p = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (p >= nent) return;
  if (!with_conj && (p & 1)) {
    if (lane == 0) out[p] = make_double2(0.0, 0.0);
    return;
  }
  double sx = 0.0, sy = 0.0;
  for (int b = lane; b < nblk; b += 64) {
    const double2 v = partial[(size_t)p * nblk + b];
    sx += v.x;
    sy += v.y;
  }
  for (int w = 32; w > 0; w >>= 1) {
    sx += __shfl_xor(sx, w, 64);
    sy += __shfl_xor(sy, w, 64);
  }
  if (lane == 0) out[p] = make_double2(sx, sy);
}

// ---------------------------------------------------------------------------
// Orbital time derivative (hprod_htot_dtorb, src/hprod/hprod_htot_dtorb.F90:2-111)
// fused with the Runge-Kutta stage update of clcrk::prop (src/prop/clcrk.cpp:55-87):
//   k   = tfac (v2 + g + h01) psi_i + sum_j psi_j X(j,i)       (i dynamical; 0 for frozen core)
//   acc += dtw * k      (when acc != nullptr)
//   next = w0 + sum_s c_s k_s  (when next != nullptr; k_s of earlier stages in kbuf)
// ---------------------------------------------------------------------------
// X(j, i) from the l-space Gram products (hprod_htot_mkxmat_fc/cc/aa1):
// fc-dyn (F + G + V), core-core and active-active (H + G + V), times xfac.
// sP: the products of column i summed from the partials ([H, G, V, F][j]),
// when lnblk > 0.
__device__ double2 x_entry(const DtorbArgs& a, int j, int i, const double2* sP) {
  const double2 z = make_double2(0.0, 0.0);
  if (i < a.nfcore || a.mval[j] != a.mval[i]) return z;
  double2 H, G, V, F;
  if (a.lnblk > 0) {
    H = sP[j];
    G = sP[a.nfun + j];
    V = sP[2 * a.nfun + j];
    F = sP[3 * a.nfun + j];
  } else {
    const int e = j * a.nfun + i;
    const int qh = a.mapH[e], qg = a.mapG[e], qv = a.mapV[e], qf = a.mapF[e];
    H = qh >= 0 ? a.lprod[qh] : z;
    G = qg >= 0 ? a.lprod[qg] : z;
    V = qv >= 0 ? a.lprod[qv] : z;
    F = qf >= 0 ? a.lprod[qf] : z;
  }
  if (j < a.nfcore) return cmul(cadd(F, cadd(G, V)), a.xfac);
  if ((j < a.ncore && i < a.ncore) || (j >= a.ncore && i >= a.ncore)) return cmul(cadd(V, cadd(G, H)), a.xfac);
  return z;
}

__device__ __forceinline__ void dtorb_body(const DtorbArgs& a, const Blk bk) {
  const size_t q = (size_t)bk.x * blockDim.x + threadIdx.x;
  const size_t nb = (size_t)a.nbas;
  const int f = bk.y;
  __shared__ double2 sXc[kMaxFun];     // column f of X when built here
  __shared__ double2 sP[4 * kMaxFun];  // its Gram products, summed from the partials
  if (a.xlocal) {  // block-uniform
    if (a.lnblk > 0) {
      // eight lanes per product: strided partial sums, then three exchanges
      const int sub = threadIdx.x & 7;
      for (int slot = threadIdx.x >> 3; slot < 4 * a.nfun; slot += 32) {
        const int w = slot / a.nfun, j = slot - w * a.nfun;
        const int* map = w == 0 ? a.mapH : w == 1 ? a.mapG : w == 2 ? a.mapV : a.mapF;
        const int qq = map[j * a.nfun + f];
        double2 acc = make_double2(0.0, 0.0);
        if (qq >= 0)
          for (int b = sub; b < a.lnblk; b += 8) acc = cadd(acc, a.lpart[(size_t)qq * a.lnblk + b]);
        for (int o = 1; o < 8; o <<= 1) {
          acc.x += __shfl_xor(acc.x, o, 64);
          acc.y += __shfl_xor(acc.y, o, 64);
        }
        if (sub == 0) sP[slot] = acc;
      }
      __syncthreads();
    }
    if ((int)threadIdx.x < a.nfun) sXc[threadIdx.x] = x_entry(a, threadIdx.x, f, sP);
    __syncthreads();
  }
  if (q >= nb) return;
  const size_t idx = (size_t)f * nb + q;
  double2 k = make_double2(0.0, 0.0);
  if (f >= a.nfcore) {
    const size_t di = (size_t)(f - a.nfcore) * nb + q;
    double2 h;
    if (a.isplit == 1) {
      h = cadd(a.v2orb[di], a.gorb[di]);
    } else {
      h = a.h01[idx];
      h = cadd(a.v2orb[di], cadd(a.gorb[di], h));
    }
    k = cmul(h, a.tfac);
    for (int j = 0; j < a.nfun; ++j) {
      if (a.mval[j] != a.mval[f]) continue;
      cfma(k, a.orb[(size_t)j * nb + q], a.xlocal ? sXc[j] : a.xmat[(size_t)f * a.nfun + j]);
    }
    if (a.ladapt && (a.ladapt == 1 || f < a.ncore) && (int)(q / (size_t)a.nr) != a.lval[f])
      k = make_double2(0.0, 0.0);
  }
  if (a.kout) a.kout[idx] = k;
  // the step-start state: read before acc (which it may alias) is written
  const double2 w0v = a.next ? a.w0[idx] : make_double2(0.0, 0.0);
  if (a.acc) {
    double2 w = a.acc[idx];
    if (a.w0out) a.w0out[idx] = w;
    w.x = fma(a.dtw, k.x, w.x);
    w.y = fma(a.dtw, k.y, w.y);
    if (a.mask && f >= a.nfcore) {
      const int r = (int)(q % a.nr);
      w = cscale(w, a.mask[r]);
    }
    a.acc[idx] = w;
  }
  if (a.next) {
    double2 t = w0v;
    for (int s = 0; s < a.nprev; ++s) {
      const double2 ks = a.kprev[s][idx];
      t.x = fma(a.cprev[s], ks.x, t.x);
      t.y = fma(a.cprev[s], ks.y, t.y);
    }
    t.x = fma(a.cnew, k.x, t.x);
    t.y = fma(a.cnew, k.y, t.y);
    a.next[idx] = t;
  }
}

__global__ void __launch_bounds__(256) k_dtorb(const DtorbArgs a) {
  dtorb_body(a, hw_blk());
}

// radial mask of the dynamical orbitals (wfn_mask, src/wfn/wfn_mask.F90:2-35)
__device__ __forceinline__ void mask_body(double2* __restrict__ wfn, const double* __restrict__ mask, int nr, int L1, int nfcore, int nfun, const Blk bk) {
  const size_t q = (size_t)bk.x * blockDim.x + threadIdx.x;
  const size_t nb = (size_t)nr * L1;
  const int f = bk.y;
  if (q >= nb || f < nfcore || f >= nfun) return;
  wfn[(size_t)f * nb + q] = cscale(wfn[(size_t)f * nb + q], mask[q % nr]);
}

__global__ void k_mask(double2* __restrict__ wfn, const double* __restrict__ mask, int nr, int L1, int nfcore,
                       int nfun) {
  mask_body(wfn, mask, nr, L1, nfcore, nfun, hw_blk());
}

// ---------------------------------------------------------------------------
// explicit instantiations / launch wrappers
// ---------------------------------------------------------------------------
template <int MODE, class TI, class TO>
static void transform_variant(int ks, dim3 grid, hipStream_t s, const TI* in, TO* out, const double* mats,
                              const float* matsf, const TransformBatch* batch, int nr, int K, int C, int Cp,
                              double2* copy) {
  if (ks == 1)
    hipLaunchKernelGGL((k_transform<MODE, 8, 1, TI, TO>), grid, dim3(256), 0, s, in, out, mats, matsf, batch, nr, K,
                       C, Cp, copy);
  else if (ks == 2)
    hipLaunchKernelGGL((k_transform<MODE, 8, 2, TI, TO>), grid, dim3(256), 0, s, in, out, mats, matsf, batch, nr, K,
                       C, Cp, copy);
  else
    hipLaunchKernelGGL((k_transform<MODE, 8, 4, TI, TO>), grid, dim3(256), 0, s, in, out, mats, matsf, batch, nr, K,
                       C, Cp, copy);
}

template <class TI, class TO>
constexpr int type_code() {
  return std::is_same<TI, double2>::value ? (std::is_same<TO, double2>::value ? 0 : 1)
                                          : (std::is_same<TO, float2>::value ? 2 : 3);
}

static void tvar_shape(int tvar, int& CT, int& U) {
  static const int t[6][2] = {{4, 4}, {2, 4}, {4, 8}, {8, 4}, {8, 2}, {2, 8}};
  const int v = tvar >= 0 && tvar < 6 ? tvar : 0;
  CT = t[v][0];
  U = t[v][1];
}

template <class T>
static void launch_fc_phase(const FcPhaseJob& j, hipStream_t s) {
  hipLaunchKernelGGL((k_fc_phase<T>), dim3((j.nr + 255) / 256, j.nlat), dim3(256), 0, s, j);
}

constexpr long long kMfmaMinWaves = 1200;  // wavefronts of the MFMA form below which the vector form runs

template <class TI, class TO>
void launch_transform_t(int mode, const TI* in, TO* out, const double* mats, const float* matsf,
                        const TransformBatch* batch, int nbatch, int nr, int K, int C, int Cp, double2* copy,
                        hipStream_t s, int path, const FcPhaseJob* fc, int tvar) {
  if (active_recorder()) {
    // batched form: the matrix-core kernels for every member (the same bits
    // as the member's own matrix-core launch), 256 threads
    if (fc != nullptr) return record_unsupported("the frozen-core phase job");
    TransformRec r;
    std::memset(&r, 0, sizeof r);
    r.in = in; r.out = out; r.mats = mats; r.matsf = matsf; r.batch = batch;
    r.nr = nr; r.K = K; r.C = C; r.Cp = Cp; r.copy = mode == 0 ? copy : nullptr;
    const int ct = (C + 15) / 16;
    if (mode == 0) {
      int CT, U;
      tvar_shape(tvar, CT, U);
      r.nx = (nr + 63) / 64; r.ny = (ct + CT - 1) / CT; r.ntile = r.nx * r.ny * nbatch;
      const int var[4] = {CT, U, type_code<TI, TO>(), 0};
      record_launch(kRecTransformW, var, dim3(r.ntile), 256, 0, r);
    } else {
      r.nx = (nr + 15) / 16; r.ny = (ct + 3) / 4; r.ntile = r.nx * r.ny * nbatch;
      const int var[4] = {1, type_code<TI, TO>(), 0, 0};
      record_launch(kRecTransformM, var, dim3(r.ntile), 256, 0, r);
    }
    return;
  }
  // matrix cores (16 radial rows per block, one wavefront per 16 columns)
  // when that gives enough wavefronts to fill the chip and at least three
  // column tiles share each input load; otherwise the vector form, whose
  // contraction split keeps short grids busy
  const int ct = (C + 15) / 16;
  const int wpb = std::min(4, ct);
  const bool mfma = ct >= 3 && (long long)((nr + 15) / 16) * nbatch * ct >= kMfmaMinWaves;
  if (path == 1 || (path == 0 && !mfma)) {
    const int rt = (nr + 63) / 64;
    // 8 columns per wavefront; split the contraction over 2 or 4 wavefronts
    // only while the grid is too small to fill the chip (a split re-reads
    // the input once per extra column block)
    int ks = 1;
    while (ks < 4 && rt * ((C + 8 * (4 / ks) - 1) / (8 * (4 / ks))) * nbatch < 800) ks *= 2;
    const int cols = 8 * (4 / ks);
    dim3 grid(rt, (C + cols - 1) / cols, nbatch);
    if (mode == 0) transform_variant<0>(ks, grid, s, in, out, mats, matsf, batch, nr, K, C, Cp, copy);
    else transform_variant<1>(ks, grid, s, in, out, mats, matsf, batch, nr, K, C, Cp, nullptr);
    if (fc != nullptr) launch_fc_phase<TO>(*fc, s);
    return;
  }
  const FcPhaseJob fj = fc != nullptr ? *fc : FcPhaseJob{};
  if (ct >= 2 && mode == 0) {
    // register-blocked form: 64 rows x 16 CT columns per block, CT column
    // tiles per wavefront, U contraction steps per load batch (for the pair
    // products, mode 1, the form below, which forms them once per block in
    // LDS, measured faster)
    const int nfcb = fc != nullptr ? ((fc->nr + 255) / 256) * fc->nlat : 0;
    auto go = [&](auto kern, int CT) {
      const int nx = (nr + 63) / 64, ny = (ct + CT - 1) / CT, ntile = nx * ny * nbatch;
      hipLaunchKernelGGL(kern, dim3(ntile + nfcb), dim3(256), 0, s, in, out, mats, matsf, batch, nr, K, C, Cp, copy,
                         nx, ny, ntile, fj);
    };
    switch (tvar) {
      case 1: go(k_transform_mfma_w<2, 4, TI, TO>, 2); break;
      case 2: go(k_transform_mfma_w<4, 8, TI, TO>, 4); break;
      case 3: go(k_transform_mfma_w<8, 4, TI, TO>, 8); break;
      case 4: go(k_transform_mfma_w<8, 2, TI, TO>, 8); break;
      case 5: go(k_transform_mfma_w<2, 8, TI, TO>, 2); break;
      default: go(k_transform_mfma_w<4, 4, TI, TO>, 4); break;
    }
    return;
  }
  const int nx = (nr + 15) / 16, ny = (ct + wpb - 1) / wpb, ntile = nx * ny * nbatch;
  const int nfcb = fc != nullptr ? ((fc->nr + 64 * wpb - 1) / (64 * wpb)) * fc->nlat : 0;
  if (mode == 0)
    hipLaunchKernelGGL((k_transform_mfma<0, TI, TO>), dim3(ntile + nfcb), dim3(64 * wpb), 0, s, in, out, mats, matsf,
                       batch, nr, K, C, Cp, copy, nx, ny, ntile, fj);
  else
    hipLaunchKernelGGL((k_transform_mfma<1, TI, TO>), dim3(ntile + nfcb), dim3(64 * wpb), 0, s, in, out, mats, matsf,
                       batch, nr, K, C, Cp, nullptr, nx, ny, ntile, fj);
}
template void launch_transform_t<double2, float2>(int, const double2*, float2*, const double*, const float*,
                                                  const TransformBatch*, int, int, int, int, int, double2*,
                                                  hipStream_t, int, const FcPhaseJob*, int);
template void launch_transform_t<float2, float2>(int, const float2*, float2*, const double*, const float*,
                                                 const TransformBatch*, int, int, int, int, int, double2*,
                                                 hipStream_t, int, const FcPhaseJob*, int);
template void launch_transform_t<float2, double2>(int, const float2*, double2*, const double*, const float*,
                                                  const TransformBatch*, int, int, int, int, int, double2*,
                                                  hipStream_t, int, const FcPhaseJob*, int);

void launch_transform(int mode, const double2* in, double2* out, const double* mats, const TransformBatch* batch,
                      int nbatch, int nr, int K, int C, int Cp, double2* copy, hipStream_t s, int path,
                      const FcPhaseJob* fc, int tvar) {
  launch_transform_t<double2, double2>(mode, in, out, mats, nullptr, batch, nbatch, nr, K, C, Cp, copy, s, path, fc,
                                       tvar);
}

// shape: 0 = 3 elements per block (one wavefront), 1 = 6 (128 threads),
// 2 = 12 (256 threads); lc = l values per block
template <int CN>
static void launch_onebody_el_t(const double2* orb, double2* h01, const OneBodyElArgs& a, hipStream_t s, int shape,
                                int lc) {
  const int nl = (a.L1 + lc - 1) / lc;
  if (shape == 0)
    hipLaunchKernelGGL((k_onebody_el<20, 3, CN>), dim3((a.nfe + 2) / 3, nl, a.nfun), dim3(64), 0, s, orb, h01, a, lc);
  else if (shape == 1)
    hipLaunchKernelGGL((k_onebody_el<20, 6, CN>), dim3((a.nfe + 5) / 6, nl, a.nfun), dim3(128), 0, s, orb, h01, a,
                       lc);
  else
    hipLaunchKernelGGL((k_onebody_el<20, 12, CN>), dim3((a.nfe + 11) / 12, nl, a.nfun), dim3(256), 0, s, orb, h01,
                       a, lc);
}

bool launch_onebody_el(const double2* orb, double2* h01, const OneBodyElArgs& a, hipStream_t s, int shape, int lc) {
  if (a.ndvr != 20) return false;
  lc = std::max(1, std::min(lc, a.L1));
  if (active_recorder()) {
    const int GE = shape == 0 ? 3 : (shape == 1 ? 6 : 12);
    OneBodyRec r;
    std::memset(&r, 0, sizeof r);
    r.orb = orb; r.h01 = h01; r.a = a; r.lc = lc;
    const int var[4] = {GE, a.cn, 0, 0};
    record_launch(kRecOneBody, var, dim3((a.nfe + GE - 1) / GE, (a.L1 + lc - 1) / lc, a.nfun),
                  ((GE * 21 + 1) + 63) / 64 * 64, 0, r);
    return true;
  }
  if (a.cn == 1)
    launch_onebody_el_t<1>(orb, h01, a, s, shape, lc);
  else if (a.cn == 2)
    launch_onebody_el_t<2>(orb, h01, a, s, shape, lc);
  else if (a.cn == 3)
    launch_onebody_el_t<3>(orb, h01, a, s, shape, lc);
  else
    launch_onebody_el_t<0>(orb, h01, a, s, shape, lc);
  return true;
}

void launch_onebody(const double2* orb, double2* h01, const OneBodyArgs& a, hipStream_t s) {
  if (active_recorder()) return record_unsupported("the band one-body kernel (ndvr != 20)");
  dim3 block(256);
  dim3 grid((a.nr + 255) / 256, a.L1, a.f1 - a.f0);
  hipLaunchKernelGGL(k_onebody, grid, block, 0, s, orb, h01, a);
}

// Bridge pass of the Poisson solve, one block per (pair, L), for grids
// whose radial row fits in LDS: the row b = rho * invr is read once
// (coalesced) into LDS and serves the bridge dots and the Schur right-hand
// sides; the Schur system is solved by cyclic reduction in LDS and the
// bridge values x_B, the multipole moment Q and the bridge rows of v are
// written out for k_poisson_interior.  Dynamic LDS: b[nr] (complex), then
// the cyclic-reduction coefficients sPA, sPG [pcr_steps * (nfe - 1)] each.
// Jobs are L-major and XCD-remapped.
template <int NI, int BS, class T>
__device__ __forceinline__ void poisson_bridge_lds_body(const T* __restrict__ rho, T* __restrict__ v, double2* __restrict__ xb, const PoissonArgs& a, const Blk bk) {
  extern __shared__ double2 dyn_lds[];
  const int tid = threadIdx.x;
  const int job = xcd_remap(bk.x, bk.nx);
  const int L = job / a.npair, p = job % a.npair;
  const int nb = a.nfe - 1, nr = a.nr, n = a.n;
  const int ni = NI > 0 ? NI : a.ni;
  const int K = a.pcr_steps;
  __shared__ double2 sS[kMaxBridge];
  __shared__ double2 sY[kMaxBridge];
  __shared__ double2 sRed[BS / 64];
  T* sR = reinterpret_cast<T*>(dyn_lds);  // the row of rho as given (b = rho invr is formed on use)
  double* sPA = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(dyn_lds) + (size_t)nr * sizeof(T));
  double* sPG = sPA + K * nb;
  const int pm = a.pair_m[p];
  const size_t row = ((size_t)p * a.L2 + L) * nr;
  T* vp = v + row;
  if (L < (pm < 0 ? -pm : pm)) {  // uniform over the block
    // the bridge rows here, the interior rows in k_poisson_interior
    for (int k = tid; k < nb; k += BS) vp[(k + 1) * n - 1] = cz<T>();
    return;
  }
  const T* rp = rho + row;
  const size_t lo = (size_t)L * nr;
  // the cyclic-reduction coefficients of this L
  {
    const double* pa = a.pcrA + (size_t)L * K * nb;
    const double* pg = a.pcrG + (size_t)L * K * nb;
    for (int q = tid; q < K * nb; q += BS) {
      sPA[q] = pa[q];
      sPG[q] = pg[q];
    }
  }
  // the row of rho into LDS and the multipole moment sum rho rpl0, four rows
  // per thread per trip with the loads issued together
  double2 m = make_double2(0.0, 0.0);
  for (int r0 = 0; r0 < nr; r0 += 4 * BS) {
    T x[4];
    double w0[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + u * BS + tid;
      x[u] = r < nr ? rp[r] : cz<T>();
      w0[u] = r < nr ? a.rpl0[lo + r] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + u * BS + tid;
      if (r < nr) sR[r] = x[u];
      const double2 xd = cto<double2>(x[u]);
      m.x = fma(xd.x, w0[u], m.x);
      m.y = fma(xd.y, w0[u], m.y);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    m.x += __shfl_xor(m.x, o, 64);
    m.y += __shfl_xor(m.y, o, 64);
  }
  if ((tid & 63) == 0) sRed[tid >> 6] = m;
  __syncthreads();
  // Schur right-hand side of bridge k (four lanes per bridge):
  // b_B - W_l(k+1) . b_I(k+1) - W_r(k) . b_I(k), the two elements around it;
  // each lane takes every fourth row, the quad adds the parts (a quad stays
  // together in the loop)
  for (int t4 = tid; t4 < 4 * nb; t4 += BS) {
    const int kb = t4 >> 2, sub = t4 & 3;
    const int bl = kb * n, br = (kb + 1) * n;  // first interior rows of elements k and k+1
    double2 cl = make_double2(0.0, 0.0), cr = make_double2(0.0, 0.0);
    for (int i = sub; i < ni; i += 4) {
      const double2 b0 = cto<double2>(sR[bl + i]), b1 = cto<double2>(sR[br + i]);
      const double wr = a.WrI[lo + bl + i], wl = a.WlI[lo + br + i];
      cr.x = fma(wr, b0.x, cr.x);
      cr.y = fma(wr, b0.y, cr.y);
      cl.x = fma(wl, b1.x, cl.x);
      cl.y = fma(wl, b1.y, cl.y);
    }
    for (int o = 1; o < 4; o <<= 1) {
      cl.x += __shfl_xor(cl.x, o, 64);
      cl.y += __shfl_xor(cl.y, o, 64);
      cr.x += __shfl_xor(cr.x, o, 64);
      cr.y += __shfl_xor(cr.y, o, 64);
    }
    if (sub == 0) {
      double2 sk = cscale(cto<double2>(sR[br - 1]), a.invr[lo + br - 1]);
      sk = csub(sk, cl);
      sk = csub(sk, cr);
      sY[kb] = sk;
    }
  }
  __syncthreads();
  double2 msum = make_double2(0.0, 0.0);
#pragma unroll
  for (int w = 0; w < BS / 64; ++w) msum = cadd(msum, sRed[w]);
  const double2 Q = cscale(msum, a.fac1[L]);
  // parallel cyclic reduction of the Schur system (per-L coefficients)
  double2* cur = sY;
  double2* nxt = sS;
  for (int k = 0, h = 1; k < K; ++k, h <<= 1) {
    if (tid < nb) {
      const int i = tid;
      double2 d = cur[i];
      if (i - h >= 0) {
        const double al = sPA[k * nb + i];
        d = make_double2(fma(al, cur[i - h].x, d.x), fma(al, cur[i - h].y, d.y));
      }
      if (i + h < nb) {
        const double ga = sPG[k * nb + i];
        d = make_double2(fma(ga, cur[i + h].x, d.x), fma(ga, cur[i + h].y, d.y));
      }
      nxt[i] = d;
    }
    __syncthreads();
    double2* t = cur;
    cur = nxt;
    nxt = t;
  }
  // bridge values x_B and their rows of v
  if (tid < nb) {
    const int gi = (tid + 1) * n - 1;
    const double2 x = cscale(cur[tid], a.pcrBinv[(size_t)L * nb + tid]);
    xb[((size_t)p * a.L2 + L) * nb + tid] = x;
    vp[gi] = cto<T>(make_double2((x.x + Q.x * a.rpl1[lo + gi]) * a.fac2[lo + gi],
                                 (x.y + Q.y * a.rpl1[lo + gi]) * a.fac2[lo + gi]));
  }
  if (tid == 0) a.Q[(size_t)p * a.L2 + L] = Q;
}

template <int NI, int BS, class T>
__global__ void __launch_bounds__(BS) k_poisson_bridge_lds(const T* __restrict__ rho, T* __restrict__ v,
                                                       double2* __restrict__ xb, const PoissonArgs a) {
  poisson_bridge_lds_body<NI, BS, T>(rho, v, xb, a, hw_blk());
}

// Interiors of the element Schur solve after k_poisson_bridge_lds: one
// block per (L, chunk of EPB elements, group of PG pairs), thread per
// interior row.  Each thread loads its row of G_e once and applies it to
// every pair of the group, whose rows b = rho * invr are staged in LDS:
//   x_I = G_e b_I - W_l x_left - W_r x_right,  v = (x + Q rpl1) * fac2.
constexpr int kPoissonPairGroup = 7;
template <int NI, class T>
__device__ __forceinline__ void poisson_interior_body(const T* __restrict__ rho, T* __restrict__ v, const double2* __restrict__ xb, const PoissonArgs& a, int epb, int nchunk, const Blk bk) {
  constexpr int PG = kPoissonPairGroup;
  const int ni = NI > 0 ? NI : a.ni;
  const int nb = a.nfe - 1, nr = a.nr, n = a.n;
  const int ngrp = (a.npair + PG - 1) / PG;
  // L-major, then element chunk, then pair group; XCD-remapped
  const int job = xcd_remap(bk.x, bk.nx);
  const int L = job / (nchunk * ngrp), ch = (job / ngrp) % nchunk, pg = job % ngrp;
  const int p0 = pg * PG, e0 = ch * epb;
  const int t = threadIdx.x;
  const int el = t / ni, i = t - el * ni;
  const int e = e0 + el;
  const bool row_ok = el < epb && e < a.nfe;
  const int r = e * n + i;
  const size_t lo = (size_t)L * nr;
  __shared__ double2 sb[PG][kPoissonInteriorThreads];
  __shared__ double2 sx[PG][34];  // bridges e0-1 .. e0+epb-1 (epb <= 32)
  __shared__ double2 sq[PG];
  __shared__ int svalid[PG];
  if (t < PG) {
    const int p = p0 + t;
    int ok = 0;
    if (p < a.npair) {
      const int pm = a.pair_m[p];
      ok = L >= (pm < 0 ? -pm : pm);
      sq[t] = ok ? a.Q[(size_t)p * a.L2 + L] : make_double2(0.0, 0.0);
    }
    svalid[t] = ok;
  }
  // stage the rows (all loads before the barrier)
  double iv = 0.0;
  if (row_ok) iv = a.invr[lo + r];
#pragma unroll
  for (int j = 0; j < PG; ++j) {
    const int p = p0 + j;
    if (row_ok && p < a.npair) sb[j][t] = cscale(cto<double2>(rho[((size_t)p * a.L2 + L) * nr + r]), iv);
  }
  for (int q = t; q < PG * (epb + 1); q += kPoissonInteriorThreads) {
    const int j = q / (epb + 1), k = q % (epb + 1);
    const int p = p0 + j, bidx = e0 - 1 + k;
    sx[j][k] = (p < a.npair && bidx >= 0 && bidx < nb) ? xb[((size_t)p * a.L2 + L) * nb + bidx]
                                                       : make_double2(0.0, 0.0);
  }
  double gk[NI > 0 ? NI : 1];
  double wl = 0.0, wr = 0.0, r1 = 0.0, f2 = 0.0;
  const double* ge = a.G + (((size_t)L * a.gstride) + (row_ok ? e : 0)) * ni * ni + i;
  if (row_ok) {
    if (NI > 0) {
#pragma unroll
      for (int k = 0; k < NI; ++k) gk[k] = ge[k * NI];
    }
    wl = a.Wl[lo + r];
    wr = a.Wr[lo + r];
    r1 = a.rpl1[lo + r];
    f2 = a.fac2[lo + r];
  }
  __syncthreads();
  if (!row_ok) return;
  const int eb = el * ni;  // this element's first row in the staged rows
#pragma unroll 1
  for (int j = 0; j < PG; ++j) {
    const int p = p0 + j;
    if (p >= a.npair) break;
    T* vp = v + ((size_t)p * a.L2 + L) * nr;
    if (!svalid[j]) {
      vp[r] = cz<T>();
      continue;
    }
    double2 y = make_double2(0.0, 0.0);
    if (NI > 0) {
#pragma unroll
      for (int k = 0; k < NI; ++k) {
        y.x = fma(gk[k], sb[j][eb + k].x, y.x);
        y.y = fma(gk[k], sb[j][eb + k].y, y.y);
      }
    } else {
      for (int k = 0; k < ni; ++k) {
        const double g = ge[k * ni];
        y.x = fma(g, sb[j][eb + k].x, y.x);
        y.y = fma(g, sb[j][eb + k].y, y.y);
      }
    }
    if (e >= 1) y = csub(y, cscale(sx[j][el], wl));
    if (e <= a.nfe - 2) y = csub(y, cscale(sx[j][el + 1], wr));
    const double2 Q = sq[j];
    vp[r] = cto<T>(make_double2((y.x + Q.x * r1) * f2, (y.y + Q.y * r1) * f2));
  }
}

template <int NI, class T>
__global__ void __launch_bounds__(kPoissonInteriorThreads) k_poisson_interior(const T* __restrict__ rho, T* __restrict__ v,
                                                         const double2* __restrict__ xb, const PoissonArgs a,
                                                         int epb, int nchunk) {
  poisson_interior_body<NI, T>(rho, v, xb, a, epb, nchunk, hw_blk());
}

constexpr int kPoissonBlock = 512;  // threads of the LDS bridge pass (>= nfe)

// LDS bytes of the LDS bridge pass; 0 when the row does not fit
template <class T>
size_t poisson_fused_lds(const PoissonArgs& a) {
  const size_t bytes = (size_t)a.nr * sizeof(T) + 2 * (size_t)a.pcr_steps * (a.nfe - 1) * sizeof(double);
  // with the 5 KB of static arrays a block stays within 64 KB of LDS (two or
  // more blocks per CU); the 80-element grids need 34 KB (four per CU)
  return bytes <= 56 * 1024 ? bytes : 0;
}

template <class T>
void launch_poisson(const T* rho, double2* xb, T* v, const PoissonArgs& a, hipStream_t s, int part) {
  const int jobs = a.npair * a.L2;
  if (active_recorder()) {
    const size_t lds = poisson_fused_lds<T>(a);
    if (!lds || a.ni != 19) return record_unsupported("the Poisson solve of a grid without the LDS bridge pass");
    PoissonRec r;
    std::memset(&r, 0, sizeof r);
    r.rho = rho; r.v = v; r.xb = xb; r.a = a;
    r.epb = std::min(32, kPoissonInteriorThreads / a.ni);
    r.nchunk = (a.nfe + r.epb - 1) / r.epb;
    const int var[4] = {type_code<T, T>(), 0, 0, 0};
    if (part != 2) record_launch(kRecPoissonBridge, var, dim3(jobs), kPoissonBlock, lds, r);
    if (part != 1) {
      const int ngrp = (a.npair + kPoissonPairGroup - 1) / kPoissonPairGroup;
      record_launch(kRecPoissonInterior, var, dim3(a.L2 * r.nchunk * ngrp), kPoissonInteriorThreads, 0, r);
    }
    return;
  }
  if (const size_t lds = poisson_fused_lds<T>(a)) {
    // bridge pass (rows staged in LDS), then the interiors with each row of
    // G_e shared by a group of pairs
    if (part != 2) {
      if (a.ni == 19)
        hipLaunchKernelGGL((k_poisson_bridge_lds<19, kPoissonBlock, T>), dim3(jobs), dim3(kPoissonBlock), lds, s, rho,
                           v, xb, a);
      else
        hipLaunchKernelGGL((k_poisson_bridge_lds<0, kPoissonBlock, T>), dim3(jobs), dim3(kPoissonBlock), lds, s, rho,
                           v, xb, a);
    }
    if (part == 1) return;
    const int epb = std::min(32, kPoissonInteriorThreads / a.ni);
    const int nchunk = (a.nfe + epb - 1) / epb;
    const int ngrp = (a.npair + kPoissonPairGroup - 1) / kPoissonPairGroup;
    const dim3 grid(a.L2 * nchunk * ngrp);
    if (a.ni == 19)
      hipLaunchKernelGGL((k_poisson_interior<19, T>), grid, dim3(kPoissonInteriorThreads), 0, s, rho, v, xb, a, epb,
                         nchunk);
    else
      hipLaunchKernelGGL((k_poisson_interior<0, T>), grid, dim3(kPoissonInteriorThreads), 0, s, rho, v, xb, a, epb,
                         nchunk);
    return;
  }
  if (part != 2) {
    if (a.ni == 19)
      hipLaunchKernelGGL((k_poisson_bridge<19, T>), dim3(jobs), dim3(256), 0, s, rho, v, xb, a);
    else
      hipLaunchKernelGGL((k_poisson_bridge<0, T>), dim3(jobs), dim3(256), 0, s, rho, v, xb, a);
  }
  if (part == 1) return;
  const int nt = std::min(1024, (a.npair * a.ni + 63) / 64 * 64);
  hipLaunchKernelGGL((k_poisson_elem<T>), dim3(a.nfe, a.L2), dim3(std::max(nt, 64)), 0, s, rho, v, xb, a);
}
template void launch_poisson<double2>(const double2*, double2*, double2*, const PoissonArgs&, hipStream_t, int);
template void launch_poisson<float2>(const float2*, double2*, float2*, const PoissonArgs&, hipStream_t, int);

template <class T>
void launch_mfprod(const T* orbg, const T* vg, T* gorbg, T* v2orbg, const MfArgs& a, hipStream_t s) {
  dim3 grid((a.nr + 255) / 256, a.nlat), block(256);
  const int key = a.nfcore * 10000 + a.nfun * 100 + a.nact;
  if (active_recorder()) {
    MfRec r;
    std::memset(&r, 0, sizeof r);
    r.orbg = orbg; r.vg = vg; r.gorbg = gorbg; r.v2orbg = v2orbg; r.a = a;
    if (a.nfcore == 0 && a.nact == 0 && (a.nfun <= 3 || a.nfun == 9 || a.nfun == 16)) {
      const int var[4] = {a.nfun, type_code<T, T>(), 0, 0};
      record_launch(kRecMfprodT, var, grid, 256, 0, r);
    } else if (a.nfcore == 0 && a.nact == 0) {
      const int var[4] = {type_code<T, T>(), 0, 0, 0};
      record_launch(kRecMfprod, var, grid, 256, 0, r);
    } else {
      record_unsupported("the mean-field products with frozen cores or active orbitals");
    }
    return;
  }
#define MF_CASE(NFC, NF, NA)                                                                            \
  case NFC * 10000 + NF * 100 + NA:                                                                     \
    hipLaunchKernelGGL((k_mfprod_t<NFC, NF, NA, T>), grid, block, 0, s, orbg, vg, gorbg, v2orbg, a);  \
    return;
  if (a.nact == 4 && a.nfun == a.nfcore + 4 && a.ncore == a.nfcore && a.nfcore <= 1) {
    // 32 points per block in complex64, 64 in FP64 (measured)
    const int bt = std::is_same<T, float2>::value ? 128 : 256;
    dim3 gq((a.nr + bt / 4 - 1) / (bt / 4), a.nlat);
    if (a.nfcore == 1) hipLaunchKernelGGL((k_mfprod_quad<1, T>), gq, dim3(bt), 0, s, orbg, vg, gorbg, v2orbg, a);
    else hipLaunchKernelGGL((k_mfprod_quad<0, T>), gq, dim3(bt), 0, s, orbg, vg, gorbg, v2orbg, a);
    return;
  }
  switch (key) {
    MF_CASE(0, 1, 0)
    MF_CASE(0, 2, 0)
    MF_CASE(0, 3, 0)
    MF_CASE(0, 9, 0)
    MF_CASE(0, 16, 0)
    MF_CASE(1, 5, 4)
    MF_CASE(0, 3, 2)
    MF_CASE(1, 5, 3)
    default:
      break;
  }
#undef MF_CASE
  hipLaunchKernelGGL((k_mfprod<T>), grid, block, 0, s, orbg, vg, gorbg, v2orbg, a);
}
template void launch_mfprod<double2>(const double2*, const double2*, double2*, double2*, const MfArgs&, hipStream_t);
template void launch_mfprod<float2>(const float2*, const float2*, float2*, float2*, const MfArgs&, hipStream_t);


int gram_blocks(const GramArgs& a, int maxblk) {
  // about 4 index values per thread (two trips of two), at most maxblk
  // blocks along the index
  return (int)std::max<size_t>(1, std::min<size_t>(maxblk, (a.len + 1023) / 1024));
}

void launch_gram(const GramArgs& a, double2* partial, int maxblk, double2* out, hipStream_t s) {
  const int nblk = gram_blocks(a, maxblk);
  if (active_recorder()) {
    if (out != nullptr) return record_unsupported("the summed Gram products (X built outside dtorb)");
    GramRec r;
    std::memset(&r, 0, sizeof r);
    r.a = a; r.partial = partial; r.nqblk = nblk;
    const int TA = 2, TB = a.with_conj_right ? 2 : 4;
    const int var[4] = {TA, TB, a.with_conj_right ? 1 : 0, a.f32};
    record_launch(kRecGram, var, dim3(nblk, ((a.nL + TA - 1) / TA) * ((a.nR + TB - 1) / TB)), 256, 0, r);
    return;
  }
  // register tiles of 8 complex sums per thread (16 doubles), which keep
  // the kernel at many wavefronts per SIMD
  if (a.with_conj_right) {
    constexpr int TA = 2, TB = 2;
    dim3 grid(nblk, ((a.nL + TA - 1) / TA) * ((a.nR + TB - 1) / TB));
    if (a.f32) hipLaunchKernelGGL((k_gram<TA, TB, true, float2>), grid, dim3(256), 0, s, a, partial, nblk);
    else hipLaunchKernelGGL((k_gram<TA, TB, true, double2>), grid, dim3(256), 0, s, a, partial, nblk);
  } else {
    constexpr int TA = 2, TB = 4;
    dim3 grid(nblk, ((a.nL + TA - 1) / TA) * ((a.nR + TB - 1) / TB));
    if (a.f32) hipLaunchKernelGGL((k_gram<TA, TB, false, float2>), grid, dim3(256), 0, s, a, partial, nblk);
    else hipLaunchKernelGGL((k_gram<TA, TB, false, double2>), grid, dim3(256), 0, s, a, partial, nblk);
  }
  if (out == nullptr) return;
  int nent = a.nL * a.nR * 2;
  hipLaunchKernelGGL(k_gram_final, dim3((nent + 3) / 4), dim3(256), 0, s, partial, nblk, nent, out,
                     a.with_conj_right);
}

// ---------------------------------------------------------------------------
// Dipole expectation values, hprod_dipole's Ehrenfest terms
// (src/hprod/hprod_dipole.F90:51-72): the property vectors z psi
// (hprod_zprodp), p_z psi (hprod_pzprodp) and a_z psi (hprod_azprodp) at
// zfac = 1, each coupling l and l+1, traced with the 1-RDM (hprod_trace1,
// src/hprod/hprod_trace1.F90:2-88): cores twice, frozen cores over their
// nradfc rows.  NF = 1: z only; NF = 3: z, p_z, a_z.  Thread per (r, l);
// block partial sums, then one block adds them in a fixed order.
// ---------------------------------------------------------------------------
// d psi(r) of one l row by the FEDVR first-derivative band (pmat), rows
// |j - r| <= n inside [0, nr)
__device__ __forceinline__ double2 radial_derivative(const double2* __restrict__ row, const double* __restrict__ band_p,
                                                     int nr, int n, int r) {
  const int jlo = r - n > 0 ? r - n : 0, jhi = r + n < nr - 1 ? r + n : nr - 1;
  double2 d = make_double2(0.0, 0.0);
  for (int j = jlo; j <= jhi; ++j) {
    const double pm = band_p[(size_t)(j - r + n) * nr + r];
    d.x = fma(row[j].x, pm, d.x);
    d.y = fma(row[j].y, pm, d.y);
  }
  return d;
}

// x * (-i s)
__device__ __forceinline__ double2 mul_minus_i(double2 x, double s) { return make_double2(s * x.y, -s * x.x); }

template <int NF>
__global__ void __launch_bounds__(256) k_dipole(const DipoleArgs a, double2* __restrict__ partial) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int l = blockIdx.y;
  double sum[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) sum[f] = 0.0;
  if (r < a.nr) {
    const double xr = a.xr[r];
    const size_t nr = a.nr;
    for (int i = 0; i < a.nfun; ++i) {
      const int m = a.mval[i], am = m < 0 ? -m : m;
      if (l < am) continue;
      if (i < a.nfcore && r >= a.nradfc) continue;
      const double* al = a.alph + (size_t)(m + a.mmax1) * a.L1;
      const double2* oi = a.orb + (size_t)i * a.L1 * nr;
      double2 op[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) op[f] = make_double2(0.0, 0.0);
      if (l + 1 < a.L1) {
        const double2 up = oi[(size_t)(l + 1) * nr + r];
        cfma(op[0], up, make_double2(xr * al[l], 0.0));
        if constexpr (NF == 3) {
          // p_z: (-i alph) d psi(l+1) + (-i alph (l+1) / r) psi(l+1)
          const double2 dup = radial_derivative(oi + (size_t)(l + 1) * nr, a.band_p, a.nr, a.ndvr, r);
          op[1] = cadd(op[1], cadd(mul_minus_i(dup, al[l]), mul_minus_i(up, al[l] * (l + 1.0) / xr)));
          // a_z: -alph Z / r^2
          cfma(op[2], up, make_double2(-al[l] * a.znuc / (xr * xr), 0.0));
        }
      }
      if (l - 1 >= am) {
        const double2 dn = oi[(size_t)(l - 1) * nr + r];
        cfma(op[0], dn, make_double2(xr * al[l - 1], 0.0));
        if constexpr (NF == 3) {
          // p_z: (-i alph) d psi(l-1) - (-i alph l / r) psi(l-1)
          const double2 ddn = radial_derivative(oi + (size_t)(l - 1) * nr, a.band_p, a.nr, a.ndvr, r);
          op[1] = cadd(op[1], cadd(mul_minus_i(ddn, al[l - 1]), mul_minus_i(dn, -(al[l - 1] * (double)l) / xr)));
          cfma(op[2], dn, make_double2(-al[l - 1] * a.znuc / (xr * xr), 0.0));
        }
      }
      // trace weight: 2 psi_i for a core, sum_j psi_j D(j, i) for an active
      double2 w;
      if (i < a.ncore) {
        w = cscale(oi[(size_t)l * nr + r], 2.0);
      } else {
        const int ia = i - a.ncore;
        w = make_double2(0.0, 0.0);
        for (int jb = 0; jb < a.nact; ++jb) {
          const int j = a.ncore + jb;
          if (a.mval[j] != m) continue;
          cfma(w, a.orb[((size_t)j * a.L1 + l) * nr + r], a.den1[jb + a.nact * ia]);
        }
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) sum[f] += w.x * op[f].x + w.y * op[f].y;  // Re conj(w) op
    }
  }
  __shared__ double red[NF][4];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double v = sum[f];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[f][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nblk = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      partial[(size_t)f * nblk + b] = make_double2((red[f][0] + red[f][1]) + (red[f][2] + red[f][3]), 0.0);
  }
}

// out[f] = sum of the n block partials of form f; NF = 3 adds the field
// terms of hprod_dipole: vel += A_z N (velocity gauge), acc -= E_z N
template <int NF>
__global__ void __launch_bounds__(256) k_dipole_final(const double2* __restrict__ partial, int n, const DipoleArgs a,
                                                      double* __restrict__ out) {
  __shared__ double red[4];
  for (int f = 0; f < NF; ++f) {
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) s += partial[(size_t)f * n + q].x;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = (red[0] + red[1]) + (red[2] + red[3]);
      if (NF == 3 && f == 1 && a.igauge == 1) v += a.lf[8] * a.nel;
      if (NF == 3 && f == 2) v -= a.lf[5] * a.nel;
      out[f] = v;
    }
    __syncthreads();
  }
}

int dipole_blocks(const DipoleArgs& a) { return (a.nr + 255) / 256 * a.L1; }

void launch_dipole(const DipoleArgs& a, int nforms, double2* partial, double* out, hipStream_t s) {
  dim3 grid((a.nr + 255) / 256, a.L1);
  const int nblk = (int)(grid.x * grid.y);
  if (nforms == 3) {
    hipLaunchKernelGGL(k_dipole<3>, grid, dim3(256), 0, s, a, partial);
    hipLaunchKernelGGL(k_dipole_final<3>, dim3(1), dim3(256), 0, s, partial, nblk, a, out);
  } else {
    hipLaunchKernelGGL(k_dipole<1>, grid, dim3(256), 0, s, a, partial);
    hipLaunchKernelGGL(k_dipole_final<1>, dim3(1), dim3(256), 0, s, partial, nblk, a, out);
  }
}

// ---------------------------------------------------------------------------
// hprod_dipole's frozen-core corrections (see kernels.hpp).  Print cadence:
// thread per point, the small sums in fixed order.
// ---------------------------------------------------------------------------
// The frozen cores' mean-field rows on the theta grid, thread per (r <
// nradfc, theta) (hprod_mfprod2_gfock's onFC branches,
// src/hprod/hprod_mfprod2.F90: fcx1 :158-200, dcj :282-304, dcx :343-365,
// actj :414-436, actx :489-513).
template <class A2>
__global__ void __launch_bounds__(256) k_gfock_fc(const A2* __restrict__ orbg, const A2* __restrict__ vg,
                                                  A2* __restrict__ out, const GfockFcArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int th = blockIdx.y;
  if (r >= a.nradfc) return;
  const size_t ng = (size_t)a.nlat * a.nr;
  const size_t g = (size_t)th * a.nr + r;
  const int nfc = a.nfcore, ncore = a.ncore, na = a.nact;
  A2 psi[kMaxFun], ag[kMaxFun], av[kMaxFun];
  for (int f = 0; f < a.nfun; ++f) {
    psi[f] = orbg[(size_t)f * ng + g];
    ag[f] = cz<A2>();
    av[f] = cz<A2>();
  }
  A2 d2v_dc = cz<A2>();
  A2 vact[kMaxActPairs];
  for (int p = 0; p < a.npair; ++p) {
    const A2 v = vg[(size_t)p * ng + g];
    const int j = a.pj[p], i = a.pi[p];
    if (a.ptype[p] == 1) {  // v(dyn j, frozen core i)
      if (j < ncore) {
        ag[i] = csub(ag[i], cmul(v, psi[j]));  // dcx on FC
      } else {
        // actx on FC: v2_i += wd1(j) v(j, i), wd1(j) = -1/2 sum_k den1(k, j) psi_k
        const int ja = j - ncore;
        A2 wd = cz<A2>();
        for (int ka = 0; ka < na; ++ka)
          if (a.mval[ncore + ka] == a.mval[j]) cfma(wd, psi[ncore + ka], cscale(cto<A2>(a.den1[ka + na * ja]), -0.5));
        cfma(av[i], wd, v);
      }
      continue;
    }
    if (j == i && j < ncore) d2v_dc = cadd(d2v_dc, cscale(v, 2.0));
    if (j >= ncore && i >= ncore) {
      vact[(j - ncore) * na + (i - ncore)] = v;
      if (j != i) vact[(i - ncore) * na + (j - ncore)] = cconj(v);
    }
  }
  for (int i = 0; i < nfc; ++i) cfma(ag[i], d2v_dc, psi[i]);  // dcj on FC
  if (ncore > 0 && na > 0) {
    A2 d = cz<A2>();
    for (int ja = 0; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka) cfma(d, vact[ja * na + ka], cto<A2>(a.den1[ka + na * ja]));
    for (int i = 0; i < nfc; ++i) cfma(av[i], d, psi[i]);  // actj on FC
  }
  for (int i = 0; i < nfc; ++i)
    for (int j = a.nfcore2; j < nfc; ++j)  // fcx1 on FC
      ag[i] = csub(ag[i], cmul(cto<A2>(a.vfc[((size_t)i * nfc + j) * ng + g]), psi[j]));
  for (int i = 0; i < nfc; ++i) {
    out[(size_t)i * ng + g] = ag[i];
    out[(size_t)(nfc + i) * ng + g] = av[i];
  }
}

template <class T>
void launch_gfock_fc(const T* orbg, const T* vg, T* out, const GfockFcArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_gfock_fc<T>), dim3((a.nradfc + 255) / 256, a.nlat), dim3(256), 0, s, orbg, vg, out, a);
}
template void launch_gfock_fc<double2>(const double2*, const double2*, double2*, const GfockFcArgs&, hipStream_t);
template void launch_gfock_fc<float2>(const float2*, const float2*, float2*, const GfockFcArgs&, hipStream_t);

int dipole_fc_quantities(const DipoleFcArgs& a) { return 2 + 3 * a.nfcore * a.nfun; }
int dipole_fc_blocks(const DipoleFcArgs& a) { return ((a.nradfc + 255) / 256) * a.L1; }

// z phi and p_z phi of orbital x at (r, l), zfac = 1 (hprod_zprodp,
// hprod_pzprodp as in k_dipole)
__device__ void zp_point(const DipoleFcArgs& a, int x, int r, int l, double2& z, double2& p) {
  const size_t nr = a.nr;
  const int m = a.mval[x], am = m < 0 ? -m : m;
  z = make_double2(0.0, 0.0);
  p = make_double2(0.0, 0.0);
  if (l < am) return;
  const double xr = a.xr[r];
  const double* al = a.alph + (size_t)(m + a.mmax1) * a.L1;
  const double2* ox = a.orb + (size_t)x * a.L1 * nr;
  if (l + 1 < a.L1) {
    const double2 up = ox[(size_t)(l + 1) * nr + r];
    cfma(z, up, make_double2(xr * al[l], 0.0));
    const double2 dup = radial_derivative(ox + (size_t)(l + 1) * nr, a.band_p, a.nr, a.ndvr, r);
    p = cadd(p, cadd(mul_minus_i(dup, al[l]), mul_minus_i(up, al[l] * (l + 1.0) / xr)));
  }
  if (l - 1 >= am) {
    const double2 dn = ox[(size_t)(l - 1) * nr + r];
    cfma(z, dn, make_double2(xr * al[l - 1], 0.0));
    const double2 ddn = radial_derivative(ox + (size_t)(l - 1) * nr, a.band_p, a.nr, a.ndvr, r);
    p = cadd(p, cadd(mul_minus_i(ddn, al[l - 1]), mul_minus_i(dn, -(al[l - 1] * (double)l) / xr)));
  }
}

// the generalised Fock vector f_x at (r, l)
__device__ double2 gfock_point(const DipoleFcArgs& a, int x, int r, int l) {
  const size_t q = (size_t)l * a.nr + r, nb = (size_t)a.L1 * a.nr;
  double2 f = a.h01[(size_t)x * nb + q];
  if (x < a.nfcore) {
    f = cadd(f, cadd(a.gfc[(size_t)x * nb + q], a.vfc[(size_t)x * nb + q]));
  } else {
    const int d = x - a.nfcore;
    f = cadd(f, cadd(a.gdyn[(size_t)d * nb + q], a.vdyn[(size_t)d * nb + q]));
    if (a.igauge == 1) {
      double2 z, p;
      zp_point(a, x, r, l, z, p);
      cfma(f, z, make_double2(a.lf[5], 0.0));
    }
  }
  return f;
}

__global__ void __launch_bounds__(256) k_dipole_fc(const DipoleFcArgs a, double2* __restrict__ partial) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int l = blockIdx.y;
  const int q = blockIdx.z;
  const int nf = a.nfun, nfc = a.nfcore;
  double2 s = make_double2(0.0, 0.0);
  if (r < a.nradfc) {
    const size_t nb = (size_t)a.L1 * a.nr, at = (size_t)l * a.nr + r;
    if (q < 2) {
      for (int i = 0; i < nfc; ++i) {
        double2 z, p;
        zp_point(a, i, r, l, z, p);
        cfma(s, cconj(q == 0 ? z : p), gfock_point(a, i, r, l));
      }
    } else if (q < 2 + nfc * nf) {
      const int i = (q - 2) / nf, x = (q - 2) % nf;
      if (a.mval[i] == a.mval[x]) s = cmul(cconj(a.orb[(size_t)i * nb + at]), gfock_point(a, x, r, l));
    } else {
      const int q2 = q - 2 - nfc * nf;
      const int which = q2 / (nf * nfc), x = (q2 % (nf * nfc)) / nfc, i = q2 % nfc;
      if (a.mval[i] == a.mval[x]) {
        double2 z, p;
        zp_point(a, x, r, l, z, p);
        s = cmul(cconj(which == 0 ? z : p), a.orb[(size_t)i * nb + at]);
      }
    }
  }
  __shared__ double2 red[4];
  for (int o = 32; o > 0; o >>= 1) {
    s.x += __shfl_xor(s.x, o, 64);
    s.y += __shfl_xor(s.y, o, 64);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nblk = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    partial[(size_t)q * nblk + b] = cadd(cadd(red[0], red[1]), cadd(red[2], red[3]));
  }
}

// vel(2) = -4 Im S0 - 2 Im T_z, acc(2) = -4 Im S1 - 2 Im T_p with the trace
// T = tr(D A, P_fc f) of hprod_trace1 over the projected f (hprod_projfc):
// T = 2 sum_{core x} sum_{fc i} F(i, x) S_A(x, i)
//   + sum_{act a} sum_{act j, m_j = m_a} conj(D(j, a)) sum_{fc i} F(i, a) S_A(j, i)
__global__ void __launch_bounds__(256) k_dipole_fc_final(const double2* __restrict__ partial, int nblk, int nq,
                                                         const DipoleFcArgs a, const double* __restrict__ d3,
                                                         double* __restrict__ out) {
  extern __shared__ double2 sq[];  // nq sums
  __shared__ double2 red[4];
  for (int q = 0; q < nq; ++q) {
    double2 s = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < nblk; b += 256) s = cadd(s, partial[(size_t)q * nblk + b]);
    for (int o = 32; o > 0; o >>= 1) {
      s.x += __shfl_xor(s.x, o, 64);
      s.y += __shfl_xor(s.y, o, 64);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) sq[q] = cadd(cadd(red[0], red[1]), cadd(red[2], red[3]));
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const int nf = a.nfun, nfc = a.nfcore, ncore = a.ncore, na = a.nact;
  auto F = [&](int i, int x) { return sq[2 + i * nf + x]; };
  auto S = [&](int w, int x, int i) { return sq[2 + nfc * nf + w * nf * nfc + x * nfc + i]; };
  double corr[2];
  for (int w = 0; w < 2; ++w) {
    double2 t = make_double2(0.0, 0.0);
    for (int x = 0; x < ncore; ++x)
      for (int i = 0; i < nfc; ++i)
        if (a.mval[i] == a.mval[x]) cfma(t, F(i, x), S(w, x, i));
    t = cscale(t, 2.0);
    for (int ia = 0; ia < na; ++ia) {
      const int x = ncore + ia;
      for (int ja = 0; ja < na; ++ja) {
        const int j = ncore + ja;
        if (a.mval[j] != a.mval[x]) continue;
        double2 u = make_double2(0.0, 0.0);
        for (int i = 0; i < nfc; ++i)
          if (a.mval[i] == a.mval[x]) cfma(u, F(i, x), S(w, j, i));
        cfma(t, cconj(a.den1[ja + na * ia]), u);
      }
    }
    corr[w] = -4.0 * sq[w].y - 2.0 * t.y;
  }
  out[0] = d3[0];
  out[1] = 0.0;
  out[2] = d3[0];
  out[3] = d3[1];
  out[4] = corr[0];
  out[5] = d3[1] + corr[0];
  out[6] = d3[2];
  out[7] = corr[1];
  out[8] = d3[2] + corr[1];
}

__global__ void k_dipole_expand(const double* __restrict__ d3, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  for (int f = 0; f < 3; ++f) {
    out[3 * f] = d3[f];
    out[3 * f + 1] = 0.0;
    out[3 * f + 2] = d3[f];
  }
}

void launch_dipole_expand(const double* d3, double* out9, hipStream_t s) {
  hipLaunchKernelGGL(k_dipole_expand, dim3(1), dim3(64), 0, s, d3, out9);
}

void launch_dipole_fc(const DipoleFcArgs& a, double2* partial, const double* d3, double* out9, hipStream_t s) {
  const int nq = dipole_fc_quantities(a);
  dim3 grid((a.nradfc + 255) / 256, a.L1, nq);
  hipLaunchKernelGGL(k_dipole_fc, grid, dim3(256), 0, s, a, partial);
  hipLaunchKernelGGL(k_dipole_fc_final, dim3(1), dim3(256), (size_t)nq * sizeof(double2), s, partial,
                     (int)(grid.x * grid.y), nq, a, d3, out9);
}

// ---------------------------------------------------------------------------
// Windowed orbital overlap (hprod_mkovlp, src/hprod/hprod_mkovlp.F90:2-91):
// ovlp[j + nf i] = sum_{l >= |m|} sum_{r < cut} conj(orb_j(r, l)) orb_i(r, l)
// for m_i == m_j, else 0.  One block per (j, i); threads stride over (l, r);
// fixed-order block sum.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_window_ovlp(const double2* __restrict__ orb, const int* __restrict__ mval,
                                                     int nf, int nr, int L1, int cut, double2* __restrict__ ovlp) {
  const int i = blockIdx.x / nf, j = blockIdx.x % nf;
  const int m = mval[i], am = m < 0 ? -m : m;
  const bool same = mval[j] == m;
  double2 acc = make_double2(0.0, 0.0);
  if (same) {
    const double2* oi = orb + (size_t)i * L1 * nr;
    const double2* oj = orb + (size_t)j * L1 * nr;
    const int n = (L1 - am) * cut;
    for (int q = threadIdx.x; q < n; q += 256) {
      const int l = am + q / cut, r = q % cut;
      const size_t x = (size_t)l * nr + r;
      acc = cadd(acc, cmulc(oj[x], oi[x]));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc.x += __shfl_xor(acc.x, o, 64);
    acc.y += __shfl_xor(acc.y, o, 64);
  }
  __shared__ double2 red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) ovlp[j + (size_t)nf * i] = cadd(cadd(red[0], red[1]), cadd(red[2], red[3]));
}

void launch_window_ovlp(const double2* orb, const int* mval, int nf, int nr, int L1, int cut, double2* ovlp,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_window_ovlp, dim3(nf * nf), dim3(256), 0, s, orb, mval, nf, nr, L1, cut, ovlp);
}

// ---------------------------------------------------------------------------
// clwfn::orth, orth_type "schmidt" (src/wfn/clwfn.cpp:313-321;
// src/wfn/wfn_orth.F90:2-176): the CI vector normalised, then each dynamical
// orbital i in order made orthogonal to every earlier orbital of the same m
// by modified Gram-Schmidt and normalised, over l >= |m| (a contiguous
// segment of the [l][r] layout).  One workgroup: every thread owns the same
// elements in every pass, so the only cross-thread traffic is the block sums.
// ---------------------------------------------------------------------------
constexpr int kOrthThreads = 1024;

__device__ double2 orth_block_sum(double2 v, double2* red, double2* bcast) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = red[0];
    for (int w = 1; w < kOrthThreads / 64; ++w) t = cadd(t, red[w]);
    *bcast = t;
  }
  __syncthreads();
  return *bcast;
}

__global__ void __launch_bounds__(kOrthThreads) k_orth(double2* __restrict__ orb, double2* __restrict__ cic,
                                                       const int* __restrict__ mval, int nr, int L1, int nfcore,
                                                       int nfun, int lcic) {
  __shared__ double2 red[kOrthThreads / 64];
  __shared__ double2 bcast;
  if (lcic > 0) {
    double2 a = make_double2(0.0, 0.0);
    for (int q = threadIdx.x; q < lcic; q += kOrthThreads) a.x += cic[q].x * cic[q].x + cic[q].y * cic[q].y;
    const double s = 1.0 / sqrt(orth_block_sum(a, red, &bcast).x);
    for (int q = threadIdx.x; q < lcic; q += kOrthThreads) cic[q] = cscale(cic[q], s);
  }
  for (int i = nfcore; i < nfun; ++i) {
    const int m = mval[i], am = m < 0 ? -m : m;
    const int n = (L1 - am) * nr;
    double2* oi = orb + ((size_t)i * L1 + am) * nr;
    for (int j = 0; j < i; ++j) {
      if (mval[j] != m) continue;
      const double2* oj = orb + ((size_t)j * L1 + am) * nr;
      double2 a = make_double2(0.0, 0.0);
      for (int q = threadIdx.x; q < n; q += kOrthThreads) a = cadd(a, cmulc(oj[q], oi[q]));
      const double2 t = orth_block_sum(a, red, &bcast);
      for (int q = threadIdx.x; q < n; q += kOrthThreads) oi[q] = make_double2(oi[q].x - (oj[q].x * t.x - oj[q].y * t.y),
                                                                              oi[q].y - (oj[q].x * t.y + oj[q].y * t.x));
    }
    double2 a = make_double2(0.0, 0.0);
    for (int q = threadIdx.x; q < n; q += kOrthThreads) a.x += oi[q].x * oi[q].x + oi[q].y * oi[q].y;
    const double s = 1.0 / sqrt(orth_block_sum(a, red, &bcast).x);
    for (int q = threadIdx.x; q < n; q += kOrthThreads) oi[q] = cscale(oi[q], s);
  }
}

void launch_orth(double2* orb, double2* cic, const int* mval, int nr, int L1, int nfcore, int nfun, int lcic,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_orth, dim3(1), dim3(kOrthThreads), 0, s, orb, cic, mval, nr, L1, nfcore, nfun, lcic);
}

// ---------------------------------------------------------------------------
// Small helpers that keep memset/memcpy nodes out of captured graphs.
// ---------------------------------------------------------------------------
__global__ void k_zero(double2* __restrict__ y, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    y[q] = make_double2(0.0, 0.0);
}

// Stage field values of the propagation step st->n, then st->n += 1
// (one thread; tdse_split2 evaluates clfield::get_value per stage).
// one thread per stage; the clock advances after every thread has read it
__device__ __forceinline__ void field_stages_body(const Pulse& p, PropClock* __restrict__ st, const StageNodes& nodes, double* __restrict__ lf, const Blk bk) {
  const long long n = st->n;
  const double t = st->t0 + (double)n * st->dt;
  const double dt = st->dt;
  __syncthreads();
  const int s = threadIdx.x;
  if (s < nodes.ns) pulse_lfield(p, t + nodes.c[s] * dt, lf + 9 * s);
  if (s == 0) st->n = n + 1;
}

__global__ void k_field_stages(const Pulse p, PropClock* __restrict__ st, const StageNodes nodes, double* __restrict__ lf) {
  field_stages_body(p, st, nodes, lf, hw_blk());
}

__global__ void k_clock_set(PropClock* __restrict__ st, double t0, double dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->t0 = t0;
  st->dt = dt;
  st->n = 0;
}

__global__ void k_to_f32(float2* __restrict__ y, const double2* __restrict__ x, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    y[q] = make_float2((float)x[q].x, (float)x[q].y);
}

static int grid_for(size_t n) { return (int)std::min<size_t>(2048, std::max<size_t>(1, (n + 255) / 256)); }

void launch_to_f32(float2* y, const double2* x, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(256), 0, s, y, x, n);
}

void launch_zero(double2* y, size_t n, hipStream_t s) {
  if (active_recorder()) return record_unsupported("a fill kernel");
  if (n) hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(256), 0, s, y, n);
}

// z = a x + b y (a null x or y drops its term), grid-stride; the vector
// operations of cletdrb's stages (clhprod_blas.cpp: copy, clear, scal,
// axpy, axpyz, axpbyz, xpyz, xmyz on a range of the state)
__global__ void k_zaxpby(size_t n, double2 a, const double2* __restrict__ x, double2 b, const double2* __restrict__ y,
                         double2* __restrict__ z) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (x) v = cmul(a, x[q]);
    if (y) cfma(v, b, y[q]);
    z[q] = v;
  }
}

void launch_zaxpby(size_t n, double2 a, const double2* x, double2 b, const double2* y, double2* z, hipStream_t s) {
  if (active_recorder()) return record_unsupported("a vector operation");
  if (n) hipLaunchKernelGGL(k_zaxpby, dim3(grid_for(n)), dim3(256), 0, s, n, a, x, b, y, z);
}

// The field arrays of one cletdrb step (cletdrb::prop4_orb with
// etd_time_of_h = 0: Field.get_value at t, t + dt/2, t + dt, Field.get_der at
// t, and the differences Field.get_value(t1, t, 1, -1)) for step st->n, one
// thread each; the clock advances after every thread has read it.
__global__ void k_etd_fields(const Pulse p, PropClock* __restrict__ st, double* __restrict__ lf) {
  const long long n = st->n;
  const double dt = st->dt;
  const double t = st->t0 + (double)n * dt;
  __syncthreads();
  const int q = threadIdx.x;
  if (q == 0) pulse_lfield(p, t, lf);
  if (q == 1) pulse_lfield(p, t + 0.5 * dt, lf + 9);
  if (q == 2) pulse_lfield(p, t + dt, lf + 18);
  if (q == 3) pulse_dlfield(p, t, lf + 27);
  if (q == 4 || q == 5) {
    double a[9], b[9];
    pulse_lfield(p, q == 4 ? t + 0.5 * dt : t + dt, a);
    pulse_lfield(p, t, b);
    for (int i = 0; i < 9; ++i) lf[9 * q + i] = 1.0 * a[i] + (-1.0) * b[i];
  }
  if (q == 0) st->n = n + 1;
}

// The field arrays of one clexpab step: lfield(t), then lfield(timeP[k]) -
// lfield(t) for the history times timeP[1..3] (prop_general,
// src/prop/clexpab.cpp:400-405), one thread each.
__global__ void k_xab_fields(const Pulse p, double t, double t1, double t2, double t3, double* __restrict__ lf) {
  const int q = threadIdx.x;
  if (q >= 4) return;
  double a[9], b[9];
  pulse_lfield(p, t, b);
  if (q == 0) {
    for (int i = 0; i < 9; ++i) lf[i] = b[i];
    return;
  }
  pulse_lfield(p, q == 1 ? t1 : (q == 2 ? t2 : t3), a);
  for (int i = 0; i < 9; ++i) lf[9 * q + i] = a[i] - b[i];
}

void launch_xab_fields(const Pulse& p, double t, const double* tp, double* lf, hipStream_t s) {
  hipLaunchKernelGGL(k_xab_fields, dim3(1), dim3(64), 0, s, p, t, tp[1], tp[2], tp[3], lf);
}

void launch_etd_fields(const Pulse& p, PropClock* st, double* lf, hipStream_t s) {
  hipLaunchKernelGGL(k_etd_fields, dim3(1), dim3(64), 0, s, p, st, lf);
}

void launch_field_stages(const Pulse& p, PropClock* st, const StageNodes& nodes, double* lf, hipStream_t s) {
  if (active_recorder()) {
    FieldRec r;
    std::memset(&r, 0, sizeof r);
    r.p = p; r.st = st; r.nodes = nodes; r.lf = lf;
    const int var[4] = {0, 0, 0, 0};
    return record_launch(kRecField, var, dim3(1), 64, 0, r);
  }
  hipLaunchKernelGGL(k_field_stages, dim3(1), dim3(64), 0, s, p, st, nodes, lf);
}

void launch_clock_set(PropClock* st, double t0, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(64), 0, s, st, t0, dt);
}

void launch_dtorb(const DtorbArgs& a, hipStream_t s) {
  if (active_recorder()) {
    const int var[4] = {0, 0, 0, 0};
    return record_launch(kRecDtorb, var, dim3((a.nbas + 255) / 256, a.nfun), 256, 0, a);
  }
  hipLaunchKernelGGL(k_dtorb, dim3((a.nbas + 255) / 256, a.nfun), dim3(256), 0, s, a);
}

void launch_mask(double2* wfn, const double* mask, int nr, int L1, int nfcore, int nfun, hipStream_t s) {
  size_t nb = (size_t)nr * L1;
  if (active_recorder()) {
    MaskRec r;
    std::memset(&r, 0, sizeof r);
    r.wfn = wfn; r.mask = mask; r.nr = nr; r.L1 = L1; r.nfcore = nfcore; r.nfun = nfun;
    const int var[4] = {0, 0, 0, 0};
    return record_launch(kRecMask, var, dim3((unsigned)((nb + 255) / 256), nfun), 256, 0, r);
  }
  hipLaunchKernelGGL(k_mask, dim3((unsigned)((nb + 255) / 256), nfun), dim3(256), 0, s, wfn, mask, nr, L1, nfcore,
                     nfun);
}

// ---------------------------------------------------------------------------
// Batched launches: one block range per member, each running the kernel
// body on the member's recorded arguments (kernels.hpp).
// ---------------------------------------------------------------------------
Recorder*& active_recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}

size_t rec_size(int kind) {
  switch (kind) {
    case kRecTransformW:
    case kRecTransformM: return sizeof(TransformRec);
    case kRecOneBody: return sizeof(OneBodyRec);
    case kRecPoissonBridge:
    case kRecPoissonInterior: return sizeof(PoissonRec);
    case kRecMfprodT:
    case kRecMfprod: return sizeof(MfRec);
    case kRecGram: return sizeof(GramRec);
    case kRecDtorb: return sizeof(DtorbArgs);
    case kRecMask: return sizeof(MaskRec);
    case kRecField: return sizeof(FieldRec);
    case kRecIslandB: return sizeof(IslandArgs);
    case kRecCnBridge:
    case kRecCnInterior:
    case kRecCnCheck: return sizeof(CnRec);
    default: return 0;
  }
}

template <int CT, int U, class TI, class TO>
__global__ void __launch_bounds__(256) k_transform_mfma_w_batch(const TransformRec* __restrict__ recs,
                                                                const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(TransformRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const TransformRec& r = stage_rec(recs, m, raw, have);
    transform_mfma_w_body<CT, U, TI, TO>(static_cast<const TI*>(r.in), static_cast<TO*>(r.out), r.mats, r.matsf,
                                         r.batch, r.nr, r.K, r.C, r.Cp, r.copy, r.nx, r.ny, r.ntile, r.fc, bk);
  });
}

template <int MODE, class TI, class TO>
__global__ void __launch_bounds__(256) k_transform_mfma_batch(const TransformRec* __restrict__ recs,
                                                              const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(TransformRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const TransformRec& r = stage_rec(recs, m, raw, have);
    transform_mfma_body<MODE, TI, TO>(static_cast<const TI*>(r.in), static_cast<TO*>(r.out), r.mats, r.matsf, r.batch,
                                      r.nr, r.K, r.C, r.Cp, r.copy, r.nx, r.ny, r.ntile, r.fc, bk);
  });
}

template <int GE, int CN>
__global__ void __launch_bounds__(((GE * 21 + 1) + 63) / 64 * 64)
    k_onebody_el_batch(const OneBodyRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(OneBodyRec)];
  int have = -1;
  batch_walk(map, [&](int m) { return CN == 2 && (gate_closed(recs[m].gate) || recs[m].a.stop[0] != 0); },
             [&](int m, const Blk& bk) {
    const OneBodyRec& r = stage_rec(recs, m, raw, have);
    onebody_el_body<20, GE, CN>(r.orb, r.h01, r.a, r.lc, bk);
  });
}

template <class T>
__global__ void __launch_bounds__(kPoissonBlock) k_poisson_bridge_lds_batch(const PoissonRec* __restrict__ recs,
                                                                            const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(PoissonRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const PoissonRec& r = stage_rec(recs, m, raw, have);
    poisson_bridge_lds_body<19, kPoissonBlock, T>(static_cast<const T*>(r.rho), static_cast<T*>(r.v), r.xb, r.a, bk);
  });
}

template <class T>
__global__ void __launch_bounds__(kPoissonInteriorThreads) k_poisson_interior_batch(const PoissonRec* __restrict__ recs,
                                                                                    const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(PoissonRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const PoissonRec& r = stage_rec(recs, m, raw, have);
    poisson_interior_body<19, T>(static_cast<const T*>(r.rho), static_cast<T*>(r.v), r.xb, r.a, r.epb, r.nchunk, bk);
  });
}

// One block per job (group g, L, element chunk): thread per interior row as
// in poisson_interior_body, with G_e's row and the row tables loaded once
// for the kPoissonSlots items of the group (the same element geometry makes
// them the same numbers as each member's own tables, so every item gets the
// bits of its member's own pass).  Jobs are ordered L-major, then chunk,
// then group, so that the groups reading one G_e share an XCD's L2.
template <class T>
__global__ void __launch_bounds__(kPoissonInteriorThreads)
    k_poisson_interior_grouped(const PoissonRec* __restrict__ recs, const int2* __restrict__ items,
                               const int* __restrict__ jobs, const PoissonShared sh) {
  constexpr int S = kPoissonSlots, NI = 19, n = 20;
  const int job = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int g = jobs[3 * job], L = jobs[3 * job + 1], ch = jobs[3 * job + 2];
  const int epb = sh.epb, e0 = ch * epb;
  const int t = threadIdx.x;
  const int el = t / NI, i = t - el * NI;
  const int e = e0 + el;
  const int r = e * n + i;
  __shared__ double2 sb[S][kPoissonInteriorThreads];
  __shared__ double2 sx[S][34];  // bridges e0-1 .. e0+epb-1 (epb <= 32)
  __shared__ double2 sq[S];
  __shared__ int sstate[S], snfe[S];  // state 0: not in range, 1: zero (L < |m|), 2: solve
  __shared__ const void* srho[S];
  __shared__ void* sv[S];
  __shared__ const double2* sxb[S];
  __shared__ size_t soff[S];  // ((p L2 + L) nr): the item's row start in rho and v
  __shared__ size_t sxoff[S];
  if (t < S) {
    const int2 it = items[g * S + t];
    int st = 0, nfe = 0;
    double2 q = make_double2(0.0, 0.0);
    if (it.x >= 0) {
      const PoissonRec& rc = recs[it.x];
      const int p = it.y, L2 = rc.a.L2;
      nfe = rc.a.nfe;
      if (L < L2 && ch < rc.nchunk) {
        const int pm = rc.a.pair_m[p];
        st = L >= (pm < 0 ? -pm : pm) ? 2 : 1;
        if (st == 2) q = rc.a.Q[(size_t)p * L2 + L];
      }
      srho[t] = rc.rho;
      sv[t] = rc.v;
      sxb[t] = rc.xb;
      soff[t] = ((size_t)p * L2 + L) * rc.a.nr;
      sxoff[t] = ((size_t)p * L2 + L) * (nfe - 1);
    }
    sstate[t] = st;
    snfe[t] = nfe;
    sq[t] = q;
  }
  __syncthreads();
  const bool in_blk = el < epb && e < sh.gstride;
  const size_t lo = (size_t)L * sh.rstride;
  double iv = 0.0;
  if (in_blk) iv = sh.invr[lo + r];
#pragma unroll
  for (int j = 0; j < S; ++j)
    if (sstate[j] == 2 && in_blk && e < snfe[j])
      sb[j][t] = cscale(cto<double2>(static_cast<const T*>(srho[j])[soff[j] + r]), iv);
  for (int q = t; q < S * (epb + 1); q += kPoissonInteriorThreads) {
    const int j = q / (epb + 1), k = q % (epb + 1);
    const int bidx = e0 - 1 + k;
    sx[j][k] = (sstate[j] == 2 && bidx >= 0 && bidx < snfe[j] - 1) ? sxb[j][sxoff[j] + bidx]
                                                                   : make_double2(0.0, 0.0);
  }
  double gk[NI];
  double wl = 0.0, wr = 0.0, r1 = 0.0, f2 = 0.0;
  if (in_blk) {
    const double* ge = sh.G + (((size_t)L * sh.gstride) + e) * NI * NI + i;
#pragma unroll
    for (int k = 0; k < NI; ++k) gk[k] = ge[k * NI];
    wl = sh.Wl[lo + r];
    wr = sh.Wr[lo + r];
    r1 = sh.rpl1[lo + r];
    f2 = sh.fac2[lo + r];
  }
  __syncthreads();
  if (!in_blk) return;
  const int eb = el * NI;
#pragma unroll 1
  for (int j = 0; j < S; ++j) {
    const int st = sstate[j], nfe = snfe[j];
    if (st == 0 || e >= nfe) continue;
    T* vp = static_cast<T*>(sv[j]) + soff[j];
    if (st == 1) {
      vp[r] = cz<T>();
      continue;
    }
    double2 y = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < NI; ++k) {
      y.x = fma(gk[k], sb[j][eb + k].x, y.x);
      y.y = fma(gk[k], sb[j][eb + k].y, y.y);
    }
    if (e >= 1) y = csub(y, cscale(sx[j][el], wl));
    if (e <= nfe - 2) y = csub(y, cscale(sx[j][el + 1], wr));
    const double2 Q = sq[j];
    vp[r] = cto<T>(make_double2((y.x + Q.x * r1) * f2, (y.y + Q.y * r1) * f2));
  }
}

bool launch_poisson_interior_grouped(int type, const void* recs, const int2* items, const int* jobs, int njobs,
                                     const PoissonShared& sh, hipStream_t s) {
  if (njobs <= 0) return true;
  const auto* rr = static_cast<const PoissonRec*>(recs);
  if (type == 0)
    hipLaunchKernelGGL(k_poisson_interior_grouped<double2>, dim3(njobs), dim3(kPoissonInteriorThreads), 0, s, rr,
                       items, jobs, sh);
  else if (type == 2)
    hipLaunchKernelGGL(k_poisson_interior_grouped<float2>, dim3(njobs), dim3(kPoissonInteriorThreads), 0, s, rr,
                       items, jobs, sh);
  else
    return false;
  return true;
}

template <int NF, class T>
__global__ void __launch_bounds__(256) k_mfprod_t_batch(const MfRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(MfRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const MfRec& r = stage_rec(recs, m, raw, have);
    mfprod_t_body<0, NF, 0, T>(static_cast<const T*>(r.orbg), static_cast<const T*>(r.vg), static_cast<T*>(r.gorbg),
                               static_cast<T*>(r.v2orbg), r.a, bk);
  });
}

template <class T>
__global__ void __launch_bounds__(256) k_mfprod_batch(const MfRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(MfRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const MfRec& r = stage_rec(recs, m, raw, have);
    mfprod_body<T>(static_cast<const T*>(r.orbg), static_cast<const T*>(r.vg), static_cast<T*>(r.gorbg),
                   static_cast<T*>(r.v2orbg), r.a, bk);
  });
}

template <int TA, int TB, bool CR, class T>
__global__ void __launch_bounds__(256) k_gram_batch(const GramRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(GramRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const GramRec& r = stage_rec(recs, m, raw, have);
    gram_body<TA, TB, CR, T>(r.a, r.partial, r.nqblk, bk);
  });
}

__global__ void __launch_bounds__(256) k_dtorb_batch(const DtorbArgs* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(DtorbArgs)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const DtorbArgs& r = stage_rec(recs, m, raw, have);
    dtorb_body(r, bk);
  });
}

__global__ void __launch_bounds__(256) k_mask_batch(const MaskRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(MaskRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const MaskRec& r = stage_rec(recs, m, raw, have);
    mask_body(r.wfn, r.mask, r.nr, r.L1, r.nfcore, r.nfun, bk);
  });
}

__global__ void __launch_bounds__(64) k_field_stages_batch(const FieldRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(FieldRec)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const FieldRec& r = stage_rec(recs, m, raw, have);
    field_stages_body(r.p, r.st, r.nodes, r.lf, bk);
  });
}

bool launch_batch_kernels(const LaunchRec& p, const void* recs, const BatchMap& map, int grid, size_t lds,
                          hipStream_t s, bool dry) {
  const dim3 g(grid), b(p.block);
#define TD1C_BATCH(K, R)                                                              \
  do {                                                                                \
    if (!dry) hipLaunchKernelGGL((K), g, b, lds, s, static_cast<const R*>(recs), map); \
  } while (0)
  switch (p.kind) {
    case kRecTransformW: {
      if (p.block != 256) return false;
      switch (p.var[0] * 100 + p.var[1] * 10 + p.var[2]) {
        case 240: TD1C_BATCH((k_transform_mfma_w_batch<2, 4, double2, double2>), TransformRec); return true;
        case 241: TD1C_BATCH((k_transform_mfma_w_batch<2, 4, double2, float2>), TransformRec); return true;
        case 242: TD1C_BATCH((k_transform_mfma_w_batch<2, 4, float2, float2>), TransformRec); return true;
        case 243: TD1C_BATCH((k_transform_mfma_w_batch<2, 4, float2, double2>), TransformRec); return true;
        case 440: TD1C_BATCH((k_transform_mfma_w_batch<4, 4, double2, double2>), TransformRec); return true;
        case 441: TD1C_BATCH((k_transform_mfma_w_batch<4, 4, double2, float2>), TransformRec); return true;
        case 442: TD1C_BATCH((k_transform_mfma_w_batch<4, 4, float2, float2>), TransformRec); return true;
        case 443: TD1C_BATCH((k_transform_mfma_w_batch<4, 4, float2, double2>), TransformRec); return true;
        default: return false;
      }
    }
    case kRecTransformM: {
      if (p.block != 256 || p.var[0] != 1) return false;
      if (p.var[1] == 0) TD1C_BATCH((k_transform_mfma_batch<1, double2, double2>), TransformRec);
      else if (p.var[1] == 2) TD1C_BATCH((k_transform_mfma_batch<1, float2, float2>), TransformRec);
      else return false;
      return true;
    }
    case kRecOneBody: {
      switch (p.var[0] * 10 + p.var[1]) {
        case 30: TD1C_BATCH((k_onebody_el_batch<3, 0>), OneBodyRec); return true;
        case 31: TD1C_BATCH((k_onebody_el_batch<3, 1>), OneBodyRec); return true;
        case 32: TD1C_BATCH((k_onebody_el_batch<3, 2>), OneBodyRec); return true;
        case 60: TD1C_BATCH((k_onebody_el_batch<6, 0>), OneBodyRec); return true;
        case 61: TD1C_BATCH((k_onebody_el_batch<6, 1>), OneBodyRec); return true;
        case 62: TD1C_BATCH((k_onebody_el_batch<6, 2>), OneBodyRec); return true;
        case 120: TD1C_BATCH((k_onebody_el_batch<12, 0>), OneBodyRec); return true;
        case 121: TD1C_BATCH((k_onebody_el_batch<12, 1>), OneBodyRec); return true;
        case 122: TD1C_BATCH((k_onebody_el_batch<12, 2>), OneBodyRec); return true;
        default: return false;
      }
    }
    case kRecPoissonBridge:
      if (p.var[0] == 0) TD1C_BATCH((k_poisson_bridge_lds_batch<double2>), PoissonRec);
      else if (p.var[0] == 2) TD1C_BATCH((k_poisson_bridge_lds_batch<float2>), PoissonRec);
      else return false;
      return true;
    case kRecPoissonInterior:
      if (p.var[0] == 0) TD1C_BATCH((k_poisson_interior_batch<double2>), PoissonRec);
      else if (p.var[0] == 2) TD1C_BATCH((k_poisson_interior_batch<float2>), PoissonRec);
      else return false;
      return true;
    case kRecMfprodT: {
      const bool f = p.var[1] == 2;
      if (p.var[1] != 0 && !f) return false;
      switch (p.var[0]) {
        case 1: if (f) TD1C_BATCH((k_mfprod_t_batch<1, float2>), MfRec); else TD1C_BATCH((k_mfprod_t_batch<1, double2>), MfRec); return true;
        case 2: if (f) TD1C_BATCH((k_mfprod_t_batch<2, float2>), MfRec); else TD1C_BATCH((k_mfprod_t_batch<2, double2>), MfRec); return true;
        case 3: if (f) TD1C_BATCH((k_mfprod_t_batch<3, float2>), MfRec); else TD1C_BATCH((k_mfprod_t_batch<3, double2>), MfRec); return true;
        case 9: if (f) TD1C_BATCH((k_mfprod_t_batch<9, float2>), MfRec); else TD1C_BATCH((k_mfprod_t_batch<9, double2>), MfRec); return true;
        case 16: if (f) TD1C_BATCH((k_mfprod_t_batch<16, float2>), MfRec); else TD1C_BATCH((k_mfprod_t_batch<16, double2>), MfRec); return true;
        default: return false;
      }
    }
    case kRecMfprod:
      if (p.var[0] == 0) TD1C_BATCH((k_mfprod_batch<double2>), MfRec);
      else if (p.var[0] == 2) TD1C_BATCH((k_mfprod_batch<float2>), MfRec);
      else return false;
      return true;
    case kRecGram:
      switch (p.var[0] * 1000 + p.var[1] * 100 + p.var[2] * 10 + p.var[3]) {
        case 2400: TD1C_BATCH((k_gram_batch<2, 4, false, double2>), GramRec); return true;
        case 2401: TD1C_BATCH((k_gram_batch<2, 4, false, float2>), GramRec); return true;
        case 2210: TD1C_BATCH((k_gram_batch<2, 2, true, double2>), GramRec); return true;
        case 2211: TD1C_BATCH((k_gram_batch<2, 2, true, float2>), GramRec); return true;
        default: return false;
      }
    case kRecDtorb:
      if (p.block != 256) return false;
      TD1C_BATCH(k_dtorb_batch, DtorbArgs);
      return true;
    case kRecMask:
      if (p.block != 256) return false;
      TD1C_BATCH(k_mask_batch, MaskRec);
      return true;
    case kRecField:
      if (p.block != 64) return false;
      TD1C_BATCH(k_field_stages_batch, FieldRec);
      return true;
    default: return false;
  }
#undef TD1C_BATCH
}

}  // namespace td1c
