// FP64 island of the hprod: CI reduced density matrices, regularised
// inverse of the 1-RDM, MO integrals, energies, the X matrix and the CI
// derivative.  All of it is O(nact^4 * lcic) scalar work on nact <= 16 and
// small CI spaces, so one workgroup does it without a host round trip.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace td1c {

namespace {

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zscal(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double zabs(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ double2 zero2() { return make_double2(0.0, 0.0); }

// y[full] += fac * E_pq x[full] on the alpha x beta product space
__device__ void apply_E(const IslandArgs& a, int p, int q, double2 fac, const double2* x, double2* y) {
  const int na = a.nstr_a, nb = a.nstr_b, n = a.nact;
  for (int ib = 0; ib < nb; ++ib)
    for (int ia = 0; ia < na; ++ia) {
      const double2 xv = x[ib * na + ia];
      if (xv.x == 0.0 && xv.y == 0.0) continue;
      const double2 fx = zmul(fac, xv);
      int t = a.exA[(ia * n + p) * n + q];
      if (t >= 0) {
        double s = (double)a.sgA[(ia * n + p) * n + q];
        y[ib * na + t] = zadd(y[ib * na + t], zscal(fx, s));
      }
      t = a.exB[(ib * n + p) * n + q];
      if (t >= 0) {
        double s = (double)a.sgB[(ib * n + p) * n + q];
        y[t * na + ia] = zadd(y[t * na + ia], zscal(fx, s));
      }
    }
}

// Hermitian eigen-decomposition by cyclic complex Jacobi rotations
// (stand-in for LAPACK zheev in hprod_invden_inv, src/hprod/hprod_invden.F90:62-107).
__device__ void herm_eig(int n, double2* A, double* w, double2* U) {
  for (int i = 0; i < n * n; ++i) U[i] = zero2();
  for (int i = 0; i < n; ++i) U[i + n * i] = make_double2(1.0, 0.0);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double m = zabs(A[p + n * q]);
        off += m * m;
      }
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double2 apq = A[p + n * q];
        double mag = zabs(apq);
        if (mag < 1e-300) continue;
        double app = A[p + n * p].x, aqq = A[q + n * q].x;
        double th = 0.5 * atan2(2.0 * mag, aqq - app);
        double c = cos(th), s = sin(th);
        double2 ph = zscal(apq, 1.0 / mag);
        double2 sph = zscal(ph, s), sphc = zscal(zconj(ph), s);
        for (int k = 0; k < n; ++k) {
          double2 akp = A[k + n * p], akq = A[k + n * q];
          A[k + n * p] = zsub(zscal(akp, c), zmul(sphc, akq));
          A[k + n * q] = zadd(zmul(sph, akp), zscal(akq, c));
        }
        for (int k = 0; k < n; ++k) {
          double2 apk = A[p + n * k], aqk = A[q + n * k];
          A[p + n * k] = zsub(zscal(apk, c), zmul(sph, aqk));
          A[q + n * k] = zadd(zmul(sphc, apk), zscal(aqk, c));
        }
        for (int k = 0; k < n; ++k) {
          double2 ukp = U[k + n * p], ukq = U[k + n * q];
          U[k + n * p] = zsub(zscal(ukp, c), zmul(sphc, ukq));
          U[k + n * q] = zadd(zmul(sph, ukp), zscal(ukq, c));
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = A[i + n * i].x;
}

// regularised inverse U diag(d/(d^2+thresh)) U^+
__device__ void invden_inv(int n, double thresh, const double2* a, double2* inva) {
  double2 A[kMaxAct * kMaxAct * 4], U[kMaxAct * kMaxAct * 4];
  double w[kMaxAct * 2];
  for (int i = 0; i < n * n; ++i) A[i] = a[i];
  herm_eig(n, A, w, U);
  for (int i = 0; i < n * n; ++i) inva[i] = zero2();
  for (int k = 0; k < n; ++k) {
    double d = w[k];
    double invd = thresh > 0.0 ? d / (d * d + thresh) : 1.0 / d;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        inva[i + n * j] = zadd(inva[i + n * j], zscal(zmul(U[i + n * k], zconj(U[j + n * k])), invd));
  }
}

}  // namespace

// Stage A: den1, den2 from the CI vector (ormas_mkden1/2), rden, rrden
// (hprod_invden, src/hprod/hprod_invden.F90:2-60).
__global__ void k_island_a(const IslandArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = a.nact;
  if (n == 0) return;
  const int nc = a.nstr_a * a.nstr_b;
  double2* D = a.ci_work;
  double2* C = D + (size_t)n * n * nc;
  for (int f = 0; f < nc; ++f) C[f] = a.det_of[f] >= 0 ? a.cic[a.det_of[f]] : zero2();
  for (size_t q = 0; q < (size_t)n * n * nc; ++q) D[q] = zero2();
  const double2 one = make_double2(1.0, 0.0);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) apply_E(a, p, q, one, C, D + (size_t)(p * n + q) * nc);
  // den1(a,b) = <E_ba>
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y) {
      double2 s = zero2();
      const double2* e = D + (size_t)(y * n + x) * nc;
      for (int d = 0; d < nc; ++d) s = zadd(s, zmul(zconj(C[d]), e[d]));
      a.den1[x + n * y] = s;
    }
  // den2(a,b,c,d) = (E_ab C)^+ (E_dc C) - delta_ad den1(c,b)
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q)
      for (int r = 0; r < n; ++r)
        for (int t = 0; t < n; ++t) {
          const double2* u = D + (size_t)(p * n + q) * nc;
          const double2* v = D + (size_t)(t * n + r) * nc;
          double2 s = zero2();
          for (int k = 0; k < nc; ++k) s = zadd(s, zmul(zconj(u[k]), v[k]));
          if (p == t) s = zsub(s, a.den1[r + n * q]);
          a.den2[p + n * (q + n * (r + n * t))] = s;
        }
  invden_inv(n, a.throcc1, a.den1, a.rden);
  if (a.ncore > 0) {
    if (n * 2 == a.nelact2) {
      for (int i = 0; i < n * n; ++i) a.rrden[i] = zero2();
    } else {
      double2 tmp[kMaxAct * kMaxAct * 4];
      for (int i = 0; i < n * n; ++i) tmp[i] = zscal(a.den1[i], -1.0);
      for (int i = 0; i < n; ++i) tmp[i + n * i].x += 2.0;
      invden_inv(n, a.throcc2, tmp, a.rrden);
    }
  }
}

// Stage B: int1e, int2e (hprod_mkint1_sph / hprod_mkint2_sph), energies
// (hprod_ene_dcore, hprod_ene_act), X matrix (hprod_htot_mkxmat_fc/cc/ca/aa1,
// src/hprod/hprod_htot_mkxmat.F90) and the CI derivative (ormas_xact2 with one
// ORMAS subspace, src/ormas/ormas_xact2.F90:2-102).
__global__ void k_island_b(const IslandArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nf = a.nfun, nfc = a.nfcore, ncore = a.ncore, na = a.nact;
  auto H = [&](int j, int i) { int q = a.mapH[j * nf + i]; return q >= 0 ? a.lprod[q] : zero2(); };
  auto G = [&](int j, int i) { int q = a.mapG[j * nf + i]; return q >= 0 ? a.lprod[q] : zero2(); };
  auto V = [&](int j, int i) { int q = a.mapV[j * nf + i]; return q >= 0 ? a.lprod[q] : zero2(); };
  auto F = [&](int j, int i) { int q = a.mapF[j * nf + i]; return q >= 0 ? a.lprod[q] : zero2(); };
  auto HGV = [&](int j, int i) { return zadd(V(j, i), zadd(G(j, i), H(j, i))); };
  // int1e(j,i) = <j|g + h1 + h0|i> for j <= i, Hermitian fill
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja <= ia; ++ja) {
      int i = ncore + ia, j = ncore + ja;
      a.int1e[ja + na * ia] = a.mval[i] == a.mval[j] ? zadd(G(j, i), H(j, i)) : zero2();
    }
  for (int ia = 0; ia < na; ++ia)
    for (int ja = ia + 1; ja < na; ++ja) a.int1e[ja + na * ia] = zconj(a.int1e[ia + na * ja]);
  // int2e(i,j,k,l), i >= j from the products, then the conjugate fill
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja <= ia; ++ja)
      for (int ka = 0; ka < na; ++ka)
        for (int la = 0; la < na; ++la) {
          int idx = ia + na * (ja + na * (ka + na * la));
          int q = a.map2e[idx];
          a.int2e[idx] = q >= 0 ? a.pprod[q] : zero2();
        }
  for (int ia = 0; ia < na; ++ia)
    for (int ja = ia + 1; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka)
        for (int la = 0; la < na; ++la)
          a.int2e[ia + na * (ja + na * (ka + na * la))] = zconj(a.int2e[ja + na * (ia + na * (la + na * ka))]);
  // energies
  double ene_dcore = 0.0, ene_act = 0.0;
  for (int i = nfc; i < ncore; ++i) {
    double2 t = zadd(G(i, i), zscal(H(i, i), 2.0));
    ene_dcore += t.x;
  }
  if (na > 0) {
    double2 t1 = zero2(), t2 = zero2();
    for (int ia = 0; ia < na; ++ia)
      for (int ja = 0; ja < na; ++ja) {
        t1 = zadd(t1, zmul(a.int1e[ia + na * ja], a.den1[ja + na * ia]));
        for (int ka = 0; ka < na; ++ka)
          for (int la = 0; la < na; ++la)
            t2 = zadd(t2, zmul(a.int2e[ia + na * (ja + na * (ka + na * la))],
                               a.den2[ja + na * (ia + na * (la + na * ka))]));
      }
    ene_act = t1.x + t2.x * 0.5;
  }
  a.ene[1] = ene_dcore;
  a.ene[2] = a.ene[0] + ene_dcore;
  a.ene[3] = ene_act;
  a.ene[4] = a.ene[2] + ene_act;

  const double2 tfac = a.icomp == 0 ? make_double2(a.dtime, 0.0) : make_double2(0.0, a.dtime);
  for (int q = 0; q < nf * nf; ++q) a.xmat[q] = zero2();
#define XM(j, i) a.xmat[(i) * nf + (j)]
  // fc-dyn block
  for (int i = nfc; i < nf; ++i)
    for (int j = 0; j < nfc; ++j)
      if (a.mval[j] == a.mval[i]) XM(j, i) = zmul(zadd(F(j, i), zadd(G(j, i), V(j, i))), tfac);
  // core-core block
  for (int i = nfc; i < ncore; ++i)
    for (int j = nfc; j < ncore; ++j)
      if (a.mval[j] == a.mval[i]) XM(j, i) = zmul(HGV(j, i), tfac);
  // core-active blocks
  if (na > 0 && ncore > nfc) {
    if (na * 2 == a.nelact2) {
      for (int i = nfc; i < ncore; ++i)
        for (int j = ncore; j < nf; ++j)
          if (a.mval[j] == a.mval[i]) {
            XM(j, i) = zmul(HGV(j, i), tfac);
            XM(i, j) = zmul(HGV(i, j), tfac);
          }
    } else {
      // F = <A|h|C>, <C|h|A>; B = -tfac rrden (2 F_AC - D F_CA^*) (mkxmat_ca :549-732)
      double2 FM[kMaxFun * kMaxFun];
      for (int q = 0; q < nf * nf; ++q) FM[q] = zero2();
      for (int i = nfc; i < ncore; ++i)
        for (int j = ncore; j < nf; ++j)
          if (a.mval[j] == a.mval[i]) {
            FM[i * nf + j] = HGV(j, i);
            FM[j * nf + i] = HGV(i, j);
          }
      for (int i = nfc; i < ncore; ++i)
        for (int j = ncore; j < nf; ++j) {
          XM(j, i) = zmul(FM[i * nf + j], tfac);
          XM(i, j) = zmul(FM[j * nf + i], tfac);
        }
      for (int i = nfc; i < ncore; ++i)
        for (int ja = 0; ja < na; ++ja) {
          int j = ncore + ja;
          double2 fji = zscal(FM[i * nf + j], 2.0);
          double2 t = zero2();
          for (int ka = 0; ka < na; ++ka) t = zadd(t, zmul(a.den1[ja + na * ka], zconj(FM[(ncore + ka) * nf + i])));
          FM[i * nf + j] = zsub(fji, t);
        }
      for (int i = nfc; i < ncore; ++i)
        for (int ja = 0; ja < na; ++ja) {
          int j = ncore + ja;
          double2 t = zero2();
          for (int ka = 0; ka < na; ++ka) t = zadd(t, zmul(a.rrden[ja + na * ka], FM[i * nf + ncore + ka]));
          double2 bji = zscal(zmul(tfac, t), -1.0);
          XM(j, i) = zadd(XM(j, i), bji);
          XM(i, j) = zadd(XM(i, j), zscal(zconj(bji), -1.0));
        }
    }
  }
  // active-active block within the single ORMAS subspace (mkxmat_aa1)
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja < na; ++ja) {
      int i = ncore + ia, j = ncore + ja;
      if (a.mval[i] == a.mval[j]) XM(j, i) = zmul(HGV(j, i), tfac);
    }
#undef XM
  // CI derivative: dC = fac (H_CI - E_act) C
  if (na > 0) {
    const int nc = a.nstr_a * a.nstr_b;
    double2* D = a.ci_work;
    double2* C = D + (size_t)na * na * nc;
    double2* S = C + nc;
    double2* T = S + nc;
    for (int f = 0; f < nc; ++f) {
      C[f] = a.det_of[f] >= 0 ? a.cic[a.det_of[f]] : zero2();
      S[f] = zero2();
    }
    for (size_t q = 0; q < (size_t)na * na * nc; ++q) D[q] = zero2();
    const double2 one = make_double2(1.0, 0.0);
    for (int r = 0; r < na; ++r)
      for (int s = 0; s < na; ++s) apply_E(a, r, s, one, C, D + (size_t)(r * na + s) * nc);
    for (int p = 0; p < na; ++p)
      for (int q = 0; q < na; ++q) {
        double2 hpq = a.int1e[p + na * q];
        for (int r = 0; r < na; ++r) hpq = zsub(hpq, zscal(a.int2e[p + na * (r + na * (r + na * q))], 0.5));
        const double2* e = D + (size_t)(p * na + q) * nc;
        for (int d = 0; d < nc; ++d) S[d] = zadd(S[d], zmul(hpq, e[d]));
      }
    for (int p = 0; p < na; ++p)
      for (int q = 0; q < na; ++q) {
        for (int d = 0; d < nc; ++d) T[d] = zero2();
        for (int r = 0; r < na; ++r)
          for (int s = 0; s < na; ++s) {
            double2 g = zscal(a.int2e[p + na * (q + na * (r + na * s))], 0.5);
            if (g.x == 0.0 && g.y == 0.0) continue;
            const double2* e = D + (size_t)(r * na + s) * nc;
            for (int d = 0; d < nc; ++d) T[d] = zadd(T[d], zmul(g, e[d]));
          }
        apply_E(a, p, q, one, T, S);
      }
    const double2 fac = a.icomp == 1 ? make_double2(0.0, -a.dtime) : make_double2(-a.dtime, 0.0);
    for (int f = 0; f < nc; ++f) {
      int d = a.det_of[f];
      if (d < 0) continue;
      double2 v = zsub(S[f], zscal(a.cic[d], ene_act));
      a.dcic[d] = zmul(v, fac);
    }
  } else {
    for (int d = 0; d < a.lcic; ++d) a.dcic[d] = zero2();
  }
}

void launch_island_a(const IslandArgs& a, hipStream_t s) { hipLaunchKernelGGL(k_island_a, dim3(1), dim3(64), 0, s, a); }
void launch_island_b(const IslandArgs& a, hipStream_t s) { hipLaunchKernelGGL(k_island_b, dim3(1), dim3(64), 0, s, a); }

}  // namespace td1c
