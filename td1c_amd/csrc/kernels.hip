// gfx950 kernels of the td1c orbital Hamiltonian product (FP64 path).
//
// Layouts (all complex values are double2 = (re, im), radial index fastest):
//   orbitals, l-space   : [orbital][l][r]      (td1c's wfn layout)
//   orbitals, theta-grid: [orbital][theta][r]
//   pair densities / potentials: [pair][L][r] and [pair][theta][r]
// Kernels are launched from api.hip; every launch is sized on the host from
// the same constants the kernels use for their bounds checks.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace td1c {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// ---------------------------------------------------------------------------
// Legendre transforms along the middle axis (bas_sph2ang1/ang2sph1/
// ang2sph2/sph2ang2, src/bas/*.F90): out[b][c][r] = sum_k in_b[k][r] M_b[k][c]
// with a real K x C matrix per batch entry (selected by the m value).
// MODE 0: input read directly; MODE 1: input formed on the fly as the pair
// density conj(psi_j) psi_i on the theta grid (hprod_mkrho2p fused in).
// Block: 64 radial lanes x TCOLS column groups; LDS holds a KT x C slab of M.
// ---------------------------------------------------------------------------
template <int MODE, int CPT>
__global__ void __launch_bounds__(256) k_transform(const double2* __restrict__ in,
                                                   double2* __restrict__ out,
                                                   const double* __restrict__ mats,
                                                   const TransformBatch* __restrict__ batch,
                                                   int nr, int K, int C) {
  constexpr int KT = 16;
  __shared__ double sM[KT * 256];
  const TransformBatch bt = batch[blockIdx.y];
  const int r = blockIdx.x * 64 + threadIdx.x;
  const int cg = threadIdx.y;            // column group 0..3
  const int c0 = cg * CPT;
  const double* M = mats + (size_t)bt.mat * K * C;
  const bool active = r < bt.rows;
  double2 acc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) acc[c] = make_double2(0.0, 0.0);
  const int tid = threadIdx.y * 64 + threadIdx.x;
  for (int k0 = 0; k0 < K; k0 += KT) {
    const int kn = min(KT, K - k0);
    __syncthreads();
    for (int q = tid; q < kn * C; q += 256) sM[(q / C) * 256 + (q % C)] = M[(size_t)(k0 + q / C) * C + (q % C)];
    __syncthreads();
    if (active) {
      for (int kk = 0; kk < kn; ++kk) {
        double2 x;
        if (MODE == 0) {
          x = in[(size_t)bt.in0 + (size_t)(k0 + kk) * nr + r];
        } else {
          double2 a = in[(size_t)bt.in0 + (size_t)(k0 + kk) * nr + r];
          double2 b = in[(size_t)bt.in1 + (size_t)(k0 + kk) * nr + r];
          x = cmulc(a, b);
        }
        const double* mrow = sM + kk * 256 + c0;
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
          double m = mrow[c];
          acc[c].x = fma(x.x, m, acc[c].x);
          acc[c].y = fma(x.y, m, acc[c].y);
        }
      }
    }
  }
  if (active) {
#pragma unroll
    for (int c = 0; c < CPT; ++c)
      if (c0 + c < C) out[(size_t)bt.out + (size_t)(c0 + c) * nr + r] = acc[c];
  }
}

// ---------------------------------------------------------------------------
// One-body part h01 = (T_l + h1) psi (hprod_tprodp, src/hprod/hprod_tprod.F90:152-235;
// hprod_zprodp, hprod_zprod.F90:108-135; hprod_pzprodp, hprod_pzprod.F90:114-213).
// Thread per (r, l, orbital).  The FEDVR band is walked with the transposed
// row bands band[d][r], d = j - r + n.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_onebody(const double2* __restrict__ orb, double2* __restrict__ h01,
                                                 const OneBodyArgs a) {
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
      if (l + 1 < a.L1) cfma(h, psi_l[a.nr + r], cscale(a.efield, xr * al[l]));
      if (l - 1 >= am) cfma(h, psi_l[-(ptrdiff_t)a.nr + r], cscale(a.efield, xr * al[l - 1]));
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
      double2 tmp = cscale(a.zfield, xr * al[l]);
      cfma(h, psi_l[a.nr + r], tmp);
    }
    if (l - 1 >= am) {
      double2 tmp = cscale(a.zfield, xr * al[l - 1]);
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
      double2 tmp1 = cmul(a.zfield, make_double2(0.0, -al[l]));
      double2 tmp2 = cmul(a.zfield, make_double2(0.0, -(al[l] * (l + 1.0)) / xr));
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
      double2 tmp1 = cmul(a.zfield, make_double2(0.0, -al[l - 1]));
      double2 tmp2 = cmul(a.zfield, make_double2(0.0, -(al[l - 1] * (double)l) / xr));
      cfma(h, dw, tmp1);
      cfma(h, p0[r], cscale(tmp2, -1.0));
    }
    // frozen core in the velocity gauge also carries E.z (hprod_htot.F90:123)
    if (fc) {
      if (l + 1 < a.L1) cfma(h, psi_l[a.nr + r], cscale(a.efield, xr * al[l]));
      if (l - 1 >= am) cfma(h, psi_l[-(ptrdiff_t)a.nr + r], cscale(a.efield, xr * al[l - 1]));
    }
  }
  *out = h;
}

// ---------------------------------------------------------------------------
// Radial Poisson solve per (pair, L) (hprod_mkv2mf_poissonp,
// src/hprod/hprod_mkv2mf.F90:118-182) by element Schur complement of the
// same FEDVR operator A_L = 2K + L(L+1)/r^2.
// P1: per (L, element, pair): y = G_e b_I, bridge contributions, multipole part.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_poisson_elem(const double2* __restrict__ rho, double2* __restrict__ x,
                                                     const PoissonArgs a) {
  const int e = blockIdx.x;
  const int L = blockIdx.y;
  const int ni = a.ni;
  __shared__ double sG[32 * 32];
  __shared__ double sCl[32], sCr[32];
  const double* G = a.G + (((size_t)L * a.nfe) + e) * ni * ni;
  for (int q = threadIdx.x; q < ni * ni; q += blockDim.x) sG[q] = G[q];
  for (int q = threadIdx.x; q < ni; q += blockDim.x) {
    sCl[q] = a.Cl[(((size_t)L * a.nfe) + e) * ni + q];
    sCr[q] = a.Cr[(((size_t)L * a.nfe) + e) * ni + q];
  }
  __syncthreads();
  const int base = e * a.n;  // interior global index of local point 1
  const double* invr = a.invr + (size_t)L * a.nr;
  const double* rpl0 = a.rpl0 + (size_t)L * a.nr;
  for (int p = threadIdx.x; p < a.npair; p += blockDim.x) {
    const int pm = a.pair_m[p];
    if (L < (pm < 0 ? -pm : pm)) continue;
    const double2* rp = rho + ((size_t)p * a.L2 + L) * a.nr;
    double2* xp = x + ((size_t)p * a.L2 + L) * a.nr;
    double2 b[32];
    double2 mp = make_double2(0.0, 0.0);
    for (int k = 0; k < ni; ++k) {
      double2 v = rp[base + k];
      mp.x = fma(v.x, rpl0[base + k], mp.x);
      mp.y = fma(v.y, rpl0[base + k], mp.y);
      b[k] = cscale(v, invr[base + k]);
    }
    if (e >= 1) {  // element e owns its left bridge in the multipole sum
      double2 v = rp[base - 1];
      mp.x = fma(v.x, rpl0[base - 1], mp.x);
      mp.y = fma(v.y, rpl0[base - 1], mp.y);
    }
    double2 cl = make_double2(0.0, 0.0), cr = make_double2(0.0, 0.0);
    for (int i = 0; i < ni; ++i) {
      double2 y = make_double2(0.0, 0.0);
      const double* g = sG + i * ni;
      for (int k = 0; k < ni; ++k) {
        y.x = fma(g[k], b[k].x, y.x);
        y.y = fma(g[k], b[k].y, y.y);
      }
      xp[base + i] = y;
      cl.x = fma(sCl[i], y.x, cl.x);
      cl.y = fma(sCl[i], y.y, cl.y);
      cr.x = fma(sCr[i], y.x, cr.x);
      cr.y = fma(sCr[i], y.y, cr.y);
    }
    size_t q = (((size_t)p * a.L2 + L) * a.nfe) + e;
    a.cl[q] = cl;
    a.cr[q] = cr;
    a.mpart[q] = mp;
  }
}

// P2: per (pair, L): tridiagonal Schur solve on the bridges + multipole sum.
__global__ void __launch_bounds__(64) k_poisson_bridge(const double2* __restrict__ rho, double2* __restrict__ x,
                                                       const PoissonArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.npair * a.L2) return;
  const int p = t / a.L2, L = t % a.L2;
  const int pm = a.pair_m[p];
  size_t q0 = ((size_t)p * a.L2 + L) * a.nfe;
  double2* xp = x + ((size_t)p * a.L2 + L) * a.nr;
  if (L < (pm < 0 ? -pm : pm)) {
    a.Q[(size_t)p * a.L2 + L] = make_double2(0.0, 0.0);
    return;
  }
  const double2* rp = rho + ((size_t)p * a.L2 + L) * a.nr;
  const double* invr = a.invr + (size_t)L * a.nr;
  const int nb = a.nfe - 1;
  const double* Sd = a.Sd + (size_t)L * nb;
  const double* Sl = a.Sl + (size_t)L * nb;
  const double* Su = a.Su + (size_t)L * nb;
  // forward: s_k = b(bridge k) - c_R(element k-1) - c_L(element k)
  double2 prev = make_double2(0.0, 0.0);
  for (int k = 0; k < nb; ++k) {
    int gi = (k + 1) * a.n - 1;
    double2 s = cscale(rp[gi], invr[gi]);
    s = csub(s, a.cr[q0 + k]);
    s = csub(s, a.cl[q0 + k + 1]);
    if (k > 0) s = csub(s, cscale(prev, Sl[k]));
    xp[gi] = s;  // holds the forward-eliminated value
    prev = s;
  }
  // backward
  double2 nxt = make_double2(0.0, 0.0);
  for (int k = nb - 1; k >= 0; --k) {
    int gi = (k + 1) * a.n - 1;
    double2 s = xp[gi];
    if (k < nb - 1) s = csub(s, cscale(nxt, Su[k]));
    s = cscale(s, 1.0 / Sd[k]);
    xp[gi] = s;
    nxt = s;
  }
  double2 m = make_double2(0.0, 0.0);
  for (int e = 0; e < a.nfe; ++e) m = cadd(m, a.mpart[q0 + e]);
  a.Q[(size_t)p * a.L2 + L] = cscale(m, a.fac1[L]);
}

// P3: per (pair, L, r): interior correction and potential assembly
// v = (x + Q r^(L+1) sqrt(w)) * 4pi/(2L+1)/(r sqrt(w)); reads the work
// buffer x (interior y values, bridge solutions) and writes v.
__global__ void __launch_bounds__(256) k_poisson_final(const double2* __restrict__ x, double2* __restrict__ v,
                                                       const PoissonArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int L = blockIdx.y, p = blockIdx.z;
  if (r >= a.nr) return;
  const double2* xp = x + ((size_t)p * a.L2 + L) * a.nr;
  double2* vp = v + ((size_t)p * a.L2 + L) * a.nr;
  const int pm = a.pair_m[p];
  if (L < (pm < 0 ? -pm : pm)) {
    vp[r] = make_double2(0.0, 0.0);
    return;
  }
  const int e = (r + 1) / a.n;         // element of r (a bridge belongs to the element on its right)
  const int mloc = r + 1 - e * a.n;    // local index 0..n-1
  double2 y = xp[r];
  const size_t k = (size_t)L * a.nr + r;
  if (mloc != 0) {  // interior point
    if (e >= 1) y = csub(y, cscale(xp[e * a.n - 1], a.Wl[k]));
    if (e <= a.nfe - 2) y = csub(y, cscale(xp[(e + 1) * a.n - 1], a.Wr[k]));
  }
  const double2 Q = a.Q[(size_t)p * a.L2 + L];
  const double rp1 = a.rpl1[k], f2 = a.fac2[k];
  double2 out;
  out.x = (y.x + Q.x * rp1) * f2;
  out.y = (y.y + Q.y * rp1) * f2;
  vp[r] = out;
}

// ---------------------------------------------------------------------------
// Mean-field application on the theta grid (hprod_mfprod2,
// src/hprod/hprod_mfprod2.F90: fcx1, dcj, dcx, actj, actx, act2).
// Thread per grid point (theta, r); each pair potential is read once.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mfprod(const double2* __restrict__ orbg, const double2* __restrict__ vg,
                                                double2* __restrict__ gorbg, double2* __restrict__ v2orbg,
                                                const MfArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int th = blockIdx.y;
  if (r >= a.nr) return;
  const size_t ng = (size_t)a.nlat * a.nr;
  const size_t g = (size_t)th * a.nr + r;
  double2 psi[kMaxFun];
  double2 ag[kMaxFun], av[kMaxFun];
  for (int f = 0; f < a.nfun; ++f) {
    psi[f] = orbg[(size_t)f * ng + g];
    ag[f] = make_double2(0.0, 0.0);
    av[f] = make_double2(0.0, 0.0);
  }
  double2 d2v_dc = make_double2(0.0, 0.0);
  double2 vact[kMaxActPairs];
  const int nfc = a.nfcore, ncore = a.ncore, na = a.nact;
  // actx needs wd1_j = sum_k psi_k (-den1(k,j)/2) for m_j = m_k
  for (int p = 0; p < a.npair; ++p) {
    const double2 v = vg[(size_t)p * ng + g];
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
      double2 wd = make_double2(0.0, 0.0);
      for (int ka = 0; ka < na; ++ka)
        if (a.mval[ncore + ka] == a.mval[j]) cfma(wd, psi[ncore + ka], cscale(a.den1[ka + na * ja], -0.5));
      cfma(av[i], wd, v);
    }
  }
  for (int f = nfc; f < a.nfun; ++f) cfma(ag[f], d2v_dc, psi[f]);  // dcj
  if (ncore > nfc && na > 0) {
    // actj: d2v = sum_jk v(j,k) den1(k,j) applied to dynamical core orbitals
    double2 d = make_double2(0.0, 0.0);
    for (int ja = 0; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka) cfma(d, vact[ja * na + ka], a.den1[ka + na * ja]);
    for (int f = nfc; f < ncore; ++f) cfma(av[f], d, psi[f]);
  }
  if (na > 0 && a.nelact2 >= 2) {
    // act2: t_i = sum_j psi_j d_ij, d_ij = sum_kl den2(j,i,l,k) v(k,l);
    // v2orbg_k += sum_i rden(i,k) t_i
    double2 t[kMaxAct];
    for (int q = 0; q < na; ++q) t[q] = make_double2(0.0, 0.0);
    for (int ia = 0; ia < na; ++ia)
      for (int ja = 0; ja <= ia; ++ja) {
        double2 d = make_double2(0.0, 0.0);
        for (int ka = 0; ka < na; ++ka)
          for (int la = 0; la < na; ++la) {
            if (a.mval[ncore + ia] + a.mval[ncore + ka] != a.mval[ncore + ja] + a.mval[ncore + la]) continue;
            const double2 dens = a.den2[ja + na * (ia + na * (la + na * ka))];
            if (fabs(dens.x) + fabs(dens.y) == 0.0) continue;
            if (sqrt(dens.x * dens.x + dens.y * dens.y) <= 1.0e-15) continue;
            cfma(d, dens, vact[ka * na + la]);
          }
        cfma(t[ia], psi[ncore + ja], d);
        if (ia > ja) cfma(t[ja], psi[ncore + ia], cconj(d));
      }
    for (int ia = 0; ia < na; ++ia)
      for (int ka = 0; ka < na; ++ka)
        if (a.mval[ncore + ia] == a.mval[ncore + ka]) cfma(av[ncore + ka], a.rden[ia + na * ka], t[ia]);
  }
  for (int f = nfc; f < a.nfun; ++f) {
    gorbg[(size_t)(f - nfc) * ng + g] = ag[f];
    v2orbg[(size_t)(f - nfc) * ng + g] = av[f];
  }
}

// ---------------------------------------------------------------------------
// Velocity-gauge frozen core on the theta grid (hprod_fcorbg,
// src/hprod/hprod_orbin.F90:68-105): orbg_fc = orbg0_fc * exp(-i A r cos(theta)).
// ---------------------------------------------------------------------------
__global__ void k_fc_phase(const double2* __restrict__ orbg0, double2* __restrict__ orbg, const double* __restrict__ xr,
                           const double* __restrict__ cost, double afield, int nr, int nradfc, int nlat, int nfc) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int th = blockIdx.y;
  if (r >= nr || th >= nlat) return;
  double ph = -afield * xr[r] * cost[th];
  double s, c;
  sincos(ph, &s, &c);
  for (int f = 0; f < nfc; ++f) {
    size_t q = ((size_t)f * nlat + th) * nr + r;
    orbg[q] = r < nradfc ? cmul(orbg0[q], make_double2(c, s)) : make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Gram products over a common index range, FP64, deterministic two-stage sum
// (the inner products of hprod_mkint1_sph, hprod_mkint2_sph, hprod_ene_dcore
// and hprod_htot_mkxmat_*):
//   S0[a][b] = sum_q opA(Lv_a[q]) * Rv_b[q],  S1[a][b] = sum_q opA(Lv_a[q]) * conj(Rv_b[q])
// opA = conj when conj_left.  One block stages a CH-long slice of every
// vector in LDS, so each vector element is read from memory once.
// ---------------------------------------------------------------------------
template <int CH>
__global__ void __launch_bounds__(256) k_gram(const GramArgs a, double2* __restrict__ partial) {
  extern __shared__ double2 sv[];  // (nL + nR) * CH
  const int nL = a.nL, nR = a.nR, nE = nL * nR;
  double2 acc0[4], acc1[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { acc0[k] = make_double2(0.0, 0.0); acc1[k] = make_double2(0.0, 0.0); }
  for (size_t q0 = (size_t)blockIdx.x * CH; q0 < a.len; q0 += (size_t)gridDim.x * CH) {
    __syncthreads();
    for (int t = threadIdx.x; t < (nL + nR) * CH; t += 256) {
      const int v = t / CH, c = t % CH;
      const size_t q = q0 + c;
      double2 x = make_double2(0.0, 0.0);
      if (q < a.len) x = v < nL ? a.left[v][q] : a.right[v - nL][q];
      sv[t] = x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = threadIdx.x + 256 * k;
      if (e >= nE) break;
      const int ia = e / nR, ib = e % nR;
      const double2* x = sv + ia * CH;
      const double2* y = sv + (nL + ib) * CH;
      double2 s0 = acc0[k], s1 = acc1[k];
      const double sa = a.conj_left ? -1.0 : 1.0;
      for (int c = 0; c < CH; ++c) {
        const double2 u = make_double2(x[c].x, sa * x[c].y), w = y[c];
        s0.x = fma(u.x, w.x, s0.x); s0.x = fma(-u.y, w.y, s0.x);
        s0.y = fma(u.x, w.y, s0.y); s0.y = fma(u.y, w.x, s0.y);
        if (a.with_conj_right) {
          s1.x = fma(u.x, w.x, s1.x); s1.x = fma(u.y, w.y, s1.x);
          s1.y = fma(-u.x, w.y, s1.y); s1.y = fma(u.y, w.x, s1.y);
        }
      }
      acc0[k] = s0;
      acc1[k] = s1;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = threadIdx.x + 256 * k;
    if (e >= nE) break;
    partial[((size_t)e * 2 + 0) * gridDim.x + blockIdx.x] = acc0[k];
    partial[((size_t)e * 2 + 1) * gridDim.x + blockIdx.x] = acc1[k];
  }
}

__global__ void k_gram_final(const double2* __restrict__ partial, int nblk, int nent, double2* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nent) return;
  double2 s = make_double2(0.0, 0.0);
  for (int b = 0; b < nblk; ++b) s = cadd(s, partial[(size_t)p * nblk + b]);
  out[p] = s;
}

// ---------------------------------------------------------------------------
// Orbital time derivative (hprod_htot_dtorb, src/hprod/hprod_htot_dtorb.F90:2-111)
// fused with the Runge-Kutta stage update of clcrk::prop (src/prop/clcrk.cpp:55-87):
//   k   = tfac (v2 + g + h01) psi_i + sum_j psi_j X(j,i)       (i dynamical; 0 for frozen core)
//   acc += dtw * k      (when acc != nullptr)
//   next = w0 + sum_s c_s k_s  (when next != nullptr; k_s of earlier stages in kbuf)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_dtorb(const DtorbArgs a) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nb = (size_t)a.nbas;
  const int f = blockIdx.y;
  if (q >= nb) return;
  const size_t idx = (size_t)f * nb + q;
  double2 k = make_double2(0.0, 0.0);
  if (f >= a.nfcore) {
    const size_t di = (size_t)(f - a.nfcore) * nb + q;
    double2 h = a.h01[idx];
    h = cadd(a.v2orb[di], cadd(a.gorb[di], h));
    k = cmul(h, a.tfac);
    for (int j = 0; j < a.nfun; ++j) {
      if (a.mval[j] != a.mval[f]) continue;
      cfma(k, a.orb[(size_t)j * nb + q], a.xmat[(size_t)f * a.nfun + j]);
    }
  }
  if (a.kout) a.kout[idx] = k;
  if (a.acc) {
    double2 w = a.acc[idx];
    w.x = fma(a.dtw, k.x, w.x);
    w.y = fma(a.dtw, k.y, w.y);
    if (a.mask && f >= a.nfcore) {
      const int r = (int)(q % a.nr);
      w = cscale(w, a.mask[r]);
    }
    a.acc[idx] = w;
  }
  if (a.next) {
    double2 t = a.w0[idx];
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

// radial mask of the dynamical orbitals (wfn_mask, src/wfn/wfn_mask.F90:2-35)
__global__ void k_mask(double2* __restrict__ wfn, const double* __restrict__ mask, int nr, int L1, int nfcore,
                       int nfun) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nb = (size_t)nr * L1;
  const int f = blockIdx.y;
  if (q >= nb || f < nfcore || f >= nfun) return;
  wfn[(size_t)f * nb + q] = cscale(wfn[(size_t)f * nb + q], mask[q % nr]);
}

// y = x0 + sum_s c_s k_s over the CI part (and any other small vectors)
__global__ void k_axpy_list(double2* __restrict__ y, const double2* __restrict__ x0, const AxpyArgs a, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double2 t = x0 ? x0[q] : y[q];
  for (int s = 0; s < a.nterm; ++s) {
    t.x = fma(a.c[s], a.k[s][q].x, t.x);
    t.y = fma(a.c[s], a.k[s][q].y, t.y);
  }
  y[q] = t;
}

// ---------------------------------------------------------------------------
// explicit instantiations / launch wrappers
// ---------------------------------------------------------------------------
void launch_transform(int mode, const double2* in, double2* out, const double* mats, const TransformBatch* batch,
                      int nbatch, int nr, int K, int C, hipStream_t s) {
  dim3 block(64, 4);
  dim3 grid((nr + 63) / 64, nbatch);
  // columns per thread: ceil(C / 4), instantiated for the sizes used
  int cpt = (C + 3) / 4;
  if (mode == 0) {
    if (cpt <= 4) hipLaunchKernelGGL((k_transform<0, 4>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else if (cpt <= 8) hipLaunchKernelGGL((k_transform<0, 8>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else if (cpt <= 16) hipLaunchKernelGGL((k_transform<0, 16>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else if (cpt <= 32) hipLaunchKernelGGL((k_transform<0, 32>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else hipLaunchKernelGGL((k_transform<0, 64>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
  } else {
    if (cpt <= 4) hipLaunchKernelGGL((k_transform<1, 4>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else if (cpt <= 8) hipLaunchKernelGGL((k_transform<1, 8>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else if (cpt <= 16) hipLaunchKernelGGL((k_transform<1, 16>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else if (cpt <= 32) hipLaunchKernelGGL((k_transform<1, 32>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
    else hipLaunchKernelGGL((k_transform<1, 64>), grid, block, 0, s, in, out, mats, batch, nr, K, C);
  }
}

void launch_onebody(const double2* orb, double2* h01, const OneBodyArgs& a, hipStream_t s) {
  dim3 block(256);
  dim3 grid((a.nr + 255) / 256, a.L1, a.f1 - a.f0);
  hipLaunchKernelGGL(k_onebody, grid, block, 0, s, orb, h01, a);
}

void launch_poisson(const double2* rho, double2* work, double2* v, const PoissonArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_poisson_elem, dim3(a.nfe, a.L2), dim3(64), 0, s, rho, work, a);
  int nt = a.npair * a.L2;
  hipLaunchKernelGGL(k_poisson_bridge, dim3((nt + 63) / 64), dim3(64), 0, s, rho, work, a);
  hipLaunchKernelGGL(k_poisson_final, dim3((a.nr + 255) / 256, a.L2, a.npair), dim3(256), 0, s, work, v, a);
}

void launch_mfprod(const double2* orbg, const double2* vg, double2* gorbg, double2* v2orbg, const MfArgs& a,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_mfprod, dim3((a.nr + 255) / 256, a.nlat), dim3(256), 0, s, orbg, vg, gorbg, v2orbg, a);
}

void launch_fc_phase(const double2* orbg0, double2* orbg, const double* xr, const double* cost, double afield, int nr,
                     int nradfc, int nlat, int nfc, hipStream_t s) {
  hipLaunchKernelGGL(k_fc_phase, dim3((nr + 255) / 256, nlat), dim3(256), 0, s, orbg0, orbg, xr, cost, afield, nr,
                     nradfc, nlat, nfc);
}

void launch_gram(const GramArgs& a, double2* partial, int nblk, double2* out, hipStream_t s) {
  constexpr int CH = 32;
  size_t shm = (size_t)(a.nL + a.nR) * CH * sizeof(double2);
  hipLaunchKernelGGL((k_gram<CH>), dim3(nblk), dim3(256), shm, s, a, partial);
  int nent = a.nL * a.nR * 2;
  hipLaunchKernelGGL(k_gram_final, dim3((nent + 63) / 64), dim3(64), 0, s, partial, nblk, nent, out);
}

void launch_dtorb(const DtorbArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_dtorb, dim3((a.nbas + 255) / 256, a.nfun), dim3(256), 0, s, a);
}

void launch_mask(double2* wfn, const double* mask, int nr, int L1, int nfcore, int nfun, hipStream_t s) {
  size_t nb = (size_t)nr * L1;
  hipLaunchKernelGGL(k_mask, dim3((unsigned)((nb + 255) / 256), nfun), dim3(256), 0, s, wfn, mask, nr, L1, nfcore,
                     nfun);
}

void launch_axpy_list(double2* y, const double2* x0, const AxpyArgs& a, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy_list, dim3((n + 255) / 256), dim3(256), 0, s, y, x0, a, n);
}

}  // namespace td1c
