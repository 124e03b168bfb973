// FP64 island of the hprod: CI reduced density matrices, regularised
// inverse of the 1-RDM, MO integrals, energies, the X matrix and the CI
// derivative.  One 256-thread workgroup does it without a host round trip:
// the small matrices live in LDS, loops over matrix entries / determinants
// are spread over the threads, and the n <= 8 Jacobi eigensolver runs on
// one lane out of LDS.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace td1c {

namespace {

constexpr int kT = 256;
constexpr int kN2 = kMaxAct * kMaxAct;          // 64
constexpr int kF2 = kMaxFun * kMaxFun;          // 256

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zscal(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double zabs(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ double2 zero2() { return make_double2(0.0, 0.0); }

// Where the CI work lives: LDS when it fits, else the global work buffer.
struct CiView {
  double2 *C, *D, *S, *T;       // D[n*n][nc] (D_pq = E_pq C), C[nc], S[nc], T[n*n][nc]
  const int *exA, *sgA, *exB, *sgB;
};

// Bytes of dynamic LDS a CI view needs (0: it does not fit, use global).
__host__ __device__ inline size_t ci_lds_bytes(int n, int nstr_a, int nstr_b) {
  const size_t nc = (size_t)nstr_a * nstr_b;
  const size_t z = nc * ((size_t)2 * n * n + 2) * sizeof(double2);
  const size_t t = (size_t)2 * (nstr_a + nstr_b) * n * n * sizeof(int);
  const size_t b = z + t;
  return b <= kIslandLds ? b : 0;
}

// Sets up the view; with LDS, copies the string tables in (caller syncs).
__device__ CiView ci_view(const IslandArgs& a, unsigned char* dyn) {
  const int n = a.nact, na = a.nstr_a, nb = a.nstr_b;
  const size_t nc = (size_t)na * nb, nt = (size_t)n * n;
  CiView v;
  if (a.ci_lds) {
    v.D = reinterpret_cast<double2*>(dyn);
    v.C = v.D + nt * nc;
    v.S = v.C + nc;
    v.T = v.S + nc;
    int* ti = reinterpret_cast<int*>(v.T + nt * nc);
    int* eA = ti;
    int* gA = eA + na * nt;
    int* eB = gA + na * nt;
    int* gB = eB + nb * nt;
    for (size_t q = threadIdx.x; q < na * nt; q += kT) {
      eA[q] = a.exA[q];
      gA[q] = a.sgA[q];
    }
    for (size_t q = threadIdx.x; q < nb * nt; q += kT) {
      eB[q] = a.exB[q];
      gB[q] = a.sgB[q];
    }
    v.exA = eA; v.sgA = gA; v.exB = eB; v.sgB = gB;
  } else {
    v.D = a.ci_work;
    v.C = v.D + nt * nc;
    v.S = v.C + nc;
    v.T = v.S + nc;
    v.exA = a.exA; v.sgA = a.sgA; v.exB = a.exB; v.sgB = a.sgB;
  }
  return v;
}

// (E_pq x)[f] for determinant f = (ib, it) by gathering from the source
// determinants: E_pq |s> = sg |t> exactly when E_qp |t> = sg |s>, so the
// source string is the E_qp image of the target string, alpha then beta.
__device__ __forceinline__ double2 gather_E(const IslandArgs& a, const CiView& v, int p, int q, const double2* x,
                                            int f) {
  const int ns = a.nstr_a, n = a.nact;
  const int ib = f / ns, it = f % ns;
  double2 t = zero2();
  const int ka = (it * n + q) * n + p;
  const int sa = v.exA[ka];
  if (sa >= 0) t = zadd(t, zscal(x[ib * ns + sa], (double)v.sgA[ka]));
  const int kb = (ib * n + q) * n + p;
  const int sb = v.exB[kb];
  if (sb >= 0) t = zadd(t, zscal(x[sb * ns + it], (double)v.sgB[kb]));
  return t;
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

// entry e = i + n j of the regularised inverse U diag(d/(d^2+thresh)) U^+
// (hprod_invden_inv, src/hprod/hprod_invden.F90:62-107)
__device__ double2 inv_entry(int n, double thresh, const double2* U, const double* w, int e) {
  const int i = e % n, j = e / n;
  double2 t = zero2();
  for (int k = 0; k < n; ++k) {
    const double d = w[k];
    const double invd = thresh > 0.0 ? d / (d * d + thresh) : 1.0 / d;
    t = zadd(t, zscal(zmul(U[i + n * k], zconj(U[j + n * k])), invd));
  }
  return t;
}

}  // namespace

// Stage A: den1, den2 from the CI vector (ormas_mkden1/2, den1(a,b) = <E_ba>,
// den2(a,b,c,d) = <E_ba E_dc> - delta_ad den1(c,b)), rden, rrden
// (hprod_invden, src/hprod/hprod_invden.F90:2-60).
__global__ void __launch_bounds__(kT) k_island_a(const IslandArgs a) {
  extern __shared__ __align__(16) unsigned char dyn[];
  __shared__ double2 s_d1[kN2], s_A[kN2], s_U[kN2];
  __shared__ double s_w[kMaxAct];
  const int n = a.nact;
  if (n == 0) return;
  const int nc = a.nstr_a * a.nstr_b;
  const CiView v = ci_view(a, dyn);
  double2* D = v.D;
  double2* C = v.C;
  for (int f = threadIdx.x; f < nc; f += kT) C[f] = a.det_of[f] >= 0 ? a.cic[a.det_of[f]] : zero2();
  __syncthreads();
  for (size_t e = threadIdx.x; e < (size_t)n * n * nc; e += kT) {
    const int pq = (int)(e / nc), f = (int)(e % nc);
    D[e] = gather_E(a, v, pq / n, pq % n, C, f);
  }
  __syncthreads();
  // den1(x,y) = sum conj(C) (E_yx C)
  for (int e = threadIdx.x; e < n * n; e += kT) {
    const int x = e % n, y = e / n;
    const double2* dv = D + (size_t)(y * n + x) * nc;
    double2 s = zero2();
    for (int d = 0; d < nc; ++d) s = zadd(s, zmul(zconj(C[d]), dv[d]));
    s_d1[e] = s;
    a.den1[e] = s;
    // actx weight of the mean-field kernel: -den1/2 where m matches
    const bool mm = a.mval[a.ncore + x] == a.mval[a.ncore + y];
    a.mfc[(size_t)n * n * n * n + e] = mm ? zscal(s, -0.5) : zero2();
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n * n * n; e += kT) {
    const int p = e % n, q = (e / n) % n, r = (e / (n * n)) % n, t = e / (n * n * n);
    const double2* u = D + (size_t)(p * n + q) * nc;
    const double2* w = D + (size_t)(t * n + r) * nc;
    double2 s = zero2();
    for (int k = 0; k < nc; ++k) s = zadd(s, zmul(zconj(u[k]), w[k]));
    if (p == t) s = zsub(s, s_d1[r + n * q]);
    a.den2[e] = s;
    // act2 coefficient of the mean-field kernel: den2 where the m values
    // balance and |den2| > 1e-15 (the terms hprod_mfprod2 keeps), else 0
    const int* mv = a.mval + a.ncore;
    const bool keep = mv[q] + mv[t] == mv[p] + mv[r] && s.x * s.x + s.y * s.y > 1.0e-30;
    a.mfc[e] = keep ? s : zero2();
  }
  // D goes to the global work buffer for stage B
  if (a.ci_lds)
    for (size_t q = threadIdx.x; q < (size_t)n * n * nc; q += kT) a.ci_work[q] = D[q];
  // rden = D^-1 and rrden = (2 - D)^-1, regularised: eigen-decomposition on
  // one lane, the n x n products on n*n lanes
  if (threadIdx.x == 0) {
    for (int i = 0; i < n * n; ++i) s_A[i] = s_d1[i];
    herm_eig(n, s_A, s_w, s_U);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kT) {
    const double2 r = inv_entry(n, a.throcc1, s_U, s_w, e);
    a.rden[e] = r;
    const bool mm = a.mval[a.ncore + e % n] == a.mval[a.ncore + e / n];
    a.mfc[(size_t)n * n * n * n + n * n + e] = mm ? r : zero2();
  }
  if (a.ncore > 0) {
    if (n * 2 == a.nelact2) {
      for (int e = threadIdx.x; e < n * n; e += kT) a.rrden[e] = zero2();
    } else {
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int i = 0; i < n * n; ++i) s_A[i] = zscal(s_d1[i], -1.0);
        for (int i = 0; i < n; ++i) s_A[i + n * i].x += 2.0;
        herm_eig(n, s_A, s_w, s_U);
      }
      __syncthreads();
      for (int e = threadIdx.x; e < n * n; e += kT) a.rrden[e] = inv_entry(n, a.throcc2, s_U, s_w, e);
    }
  }
}

// Stage B: int1e, int2e (hprod_mkint1_sph / hprod_mkint2_sph), energies
// (hprod_ene_dcore, hprod_ene_act), X matrix (hprod_htot_mkxmat_fc/cc/ca/aa1,
// src/hprod/hprod_htot_mkxmat.F90) and the CI derivative (ormas_xact2 with one
// ORMAS subspace, src/ormas/ormas_xact2.F90:2-102).
__device__ __forceinline__ void island_b_body(const IslandArgs& a, const Blk bk) {
  extern __shared__ __align__(16) unsigned char dyn[];
  __shared__ double2 s_H[kF2], s_G[kF2], s_V[kF2], s_F[kF2], s_X[kF2], s_FM[kF2];
  __shared__ double2 s_hp[kN2];
  __shared__ double2 s_i1[kN2], s_d1[kN2], s_rr[kN2];
  __shared__ double s_red[kT / 64];
  __shared__ double s_eact;
  __shared__ int s_mval[kMaxFun];
  __shared__ int s_q[4 * kF2];  // the four l-space product maps (partial-sum mode)
  const int nf = a.nfun, nfc = a.nfcore, ncore = a.ncore, na = a.nact;
  const int tid = threadIdx.x;
  // every global input first (the index maps, then what they point at), so
  // the gather costs two memory round trips
  double2* I2 = a.int2e;
  if (a.i2_lds) I2 = reinterpret_cast<double2*>(dyn + a.ci_lds_off);
  for (int e = tid; e < nf; e += kT) s_mval[e] = a.mval[e];
  const int nf2 = nf * nf;
  if (a.lnblk > 0) {
    // the l-space products from their partial sums: the maps first, then
    // four lanes per referenced entry (quads stay together in the loop)
    for (int e = tid; e < nf2; e += kT) {
      s_q[e] = a.mapH[e];
      s_q[nf2 + e] = a.mapG[e];
      s_q[2 * nf2 + e] = a.mapV[e];
      s_q[3 * nf2 + e] = a.mapF[e];
      s_X[e] = zero2();
    }
    __syncthreads();
    for (int slot = tid >> 2; slot < 4 * nf2; slot += kT / 4) {
      const int q = s_q[slot], sub = tid & 3;
      double2 acc = zero2();
      if (q >= 0)
        for (int b = sub; b < a.lnblk; b += 4) acc = zadd(acc, a.lpart[(size_t)q * a.lnblk + b]);
      acc.x += __shfl_xor(acc.x, 1, 64);
      acc.y += __shfl_xor(acc.y, 1, 64);
      acc.x += __shfl_xor(acc.x, 2, 64);
      acc.y += __shfl_xor(acc.y, 2, 64);
      if (sub == 0) {
        const int w = slot / nf2, e = slot - w * nf2;
        double2* dst = w == 0 ? s_H : w == 1 ? s_G : w == 2 ? s_V : s_F;
        dst[e] = acc;
      }
    }
  } else {
    for (int e = tid; e < nf2; e += kT) {
      const int qh = a.mapH[e], qg = a.mapG[e], qv = a.mapV[e], qf = a.mapF[e];
      s_H[e] = qh >= 0 ? a.lprod[qh] : zero2();
      s_G[e] = qg >= 0 ? a.lprod[qg] : zero2();
      s_V[e] = qv >= 0 ? a.lprod[qv] : zero2();
      s_F[e] = qf >= 0 ? a.lprod[qf] : zero2();
      s_X[e] = zero2();
    }
  }
  for (int e = tid; e < na * na; e += kT) {
    s_d1[e] = a.den1[e];
    s_rr[e] = a.rrden[e];
  }
  // int2e(i,j,k,l): i >= j from the products, the rest by conjugate symmetry;
  // kept in LDS (after the CI view) when it fits
  for (int e = tid; e < na * na * na * na; e += kT) {
    const int ia = e % na, ja = (e / na) % na, ka = (e / (na * na)) % na, la = e / (na * na * na);
    int src = e;
    bool cj = false;
    if (ja > ia) {
      src = ja + na * (ia + na * (la + na * ka));
      cj = true;
    }
    const int q = a.map2e[src];
    double2 v = q >= 0 ? a.pprod[q] : zero2();
    if (cj) v = zconj(v);
    I2[e] = v;
    if (a.i2_lds) a.int2e[e] = v;
  }
  // inputs of the CI derivative and of the Runge-Kutta update that do not
  // depend on the Gram products, loaded together with the gather above
  const int nc = na > 0 ? a.nstr_a * a.nstr_b : 0;
  const CiView v = ci_view(a, dyn);
  if (na > 0) {
    if (a.ci_lds)
      for (size_t q = tid; q < (size_t)na * na * nc; q += kT) v.D[q] = a.ci_work[q];
    for (int f = tid; f < nc; f += kT) v.C[f] = a.det_of[f] >= 0 ? a.cic[a.det_of[f]] : zero2();
  }
  const bool rk_pre = a.kout != nullptr && tid < a.lcic;  // this thread's first CI coefficient
  double2 rk_w0 = zero2(), rk_acc = zero2(), rk_kp[kMaxStages];
  if (rk_pre) {
    if (a.acc != nullptr) rk_acc = a.acc[tid];
    if (a.next != nullptr) {
      rk_w0 = a.w0[tid];
      for (int j = 0; j < a.nprev; ++j) rk_kp[j] = a.kprev[j][tid];
    }
  }
  __syncthreads();
#define H(j, i) s_H[(j) * nf + (i)]
#define G(j, i) s_G[(j) * nf + (i)]
#define V(j, i) s_V[(j) * nf + (i)]
#define F(j, i) s_F[(j) * nf + (i)]
#define HGV(j, i) zadd(V(j, i), zadd(G(j, i), H(j, i)))
  // int1e(j,i) = <j|g + h1 + h0|i> for j <= i, Hermitian fill
  for (int e = tid; e < na * na; e += kT) {
    const int ja = e % na, ia = e / na;
    const int i = ncore + ia, j = ncore + ja;
    double2 v = zero2();
    if (ja <= ia) {
      if (s_mval[i] == s_mval[j]) v = zadd(G(j, i), H(j, i));
    } else if (s_mval[i] == s_mval[j]) {
      v = zconj(zadd(G(i, j), H(i, j)));
    }
    s_i1[e] = v;
    a.int1e[e] = v;
  }
  __syncthreads();
  // energies: ene_act = Re sum int1e(i,j) den1(j,i) + 1/2 Re sum int2e(i,j,k,l) den2(j,i,l,k)
  double part = 0.0;
  for (int e = tid; e < na * na * na * na; e += kT) {
    const int ia = e % na, ja = (e / na) % na, ka = (e / (na * na)) % na, la = e / (na * na * na);
    part += 0.5 * zmul(I2[e], a.den2[ja + na * (ia + na * (la + na * ka))]).x;
  }
  for (int e = tid; e < na * na; e += kT) {
    const int ia = e % na, ja = e / na;
    part += zmul(s_i1[ia + na * ja], s_d1[ja + na * ia]).x;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
  if ((tid & 63) == 0) s_red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) {
    double ene_dcore = 0.0;
    for (int i = nfc; i < ncore; ++i) ene_dcore += zadd(G(i, i), zscal(H(i, i), 2.0)).x;
    const double ene_act = na > 0 ? (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]) : 0.0;
    s_eact = ene_act;
    a.ene[1] = ene_dcore;
    a.ene[2] = a.ene[0] + ene_dcore;
    a.ene[3] = ene_act;
    a.ene[4] = a.ene[2] + ene_act;
  }
  const double2 tfac = a.icomp == 0 ? make_double2(a.dtime, 0.0) : make_double2(0.0, a.dtime);
#define XM(j, i) s_X[(i) * nf + (j)]
  // fc-dyn, core-core and active-active blocks (one entry per thread)
  for (int e = tid; e < nf * nf; e += kT) {
    const int j = e % nf, i = e / nf;
    if (s_mval[j] != s_mval[i]) continue;
    if (j < nfc && i >= nfc) XM(j, i) = zmul(zadd(F(j, i), zadd(G(j, i), V(j, i))), tfac);
    else if (j >= nfc && j < ncore && i >= nfc && i < ncore) XM(j, i) = zmul(HGV(j, i), tfac);
    else if (j >= ncore && i >= ncore) XM(j, i) = zmul(HGV(j, i), tfac);
  }
  __syncthreads();
  // core-active blocks (mkxmat_ca / ca0), small: one lane
  if (tid == 0 && na > 0 && ncore > nfc) {
    if (na * 2 == a.nelact2) {
      for (int i = nfc; i < ncore; ++i)
        for (int j = ncore; j < nf; ++j)
          if (s_mval[j] == s_mval[i]) {
            XM(j, i) = zmul(HGV(j, i), tfac);
            XM(i, j) = zmul(HGV(i, j), tfac);
          }
    } else {
      double2* FM = s_FM;
      for (int q = 0; q < nf * nf; ++q) FM[q] = zero2();
      for (int i = nfc; i < ncore; ++i)
        for (int j = ncore; j < nf; ++j)
          if (s_mval[j] == s_mval[i]) {
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
          for (int ka = 0; ka < na; ++ka) t = zadd(t, zmul(s_d1[ja + na * ka], zconj(FM[(ncore + ka) * nf + i])));
          FM[i * nf + j] = zsub(fji, t);
        }
      for (int i = nfc; i < ncore; ++i)
        for (int ja = 0; ja < na; ++ja) {
          int j = ncore + ja;
          double2 t = zero2();
          for (int ka = 0; ka < na; ++ka) t = zadd(t, zmul(s_rr[ja + na * ka], FM[i * nf + ncore + ka]));
          double2 bji = zscal(zmul(tfac, t), -1.0);
          XM(j, i) = zadd(XM(j, i), bji);
          XM(i, j) = zadd(XM(i, j), zscal(zconj(bji), -1.0));
        }
    }
  }
  __syncthreads();
  for (int e = tid; e < nf * nf; e += kT) a.xmat[e] = s_X[e];
#undef XM
#undef H
#undef G
#undef V
#undef F
#undef HGV
  // CI derivative: dC = fac (H_CI - E_act) C, with D_pq = E_pq C from stage A
  if (na > 0) {
    double2* D = v.D;
    double2* C = v.C;
    double2* S = v.S;
    double2* T = v.T;
    // h'_pq = int1e(p,q) - 1/2 sum_r int2e(p,r,r,q)
    for (int e = tid; e < na * na; e += kT) {
      const int p = e % na, q = e / na;
      double2 hpq = s_i1[p + na * q];
      for (int r = 0; r < na; ++r) hpq = zsub(hpq, zscal(I2[p + na * (r + na * (r + na * q))], 0.5));
      s_hp[e] = hpq;
    }
    __syncthreads();
    // T_pq = 1/2 sum_rs int2e(p,q,r,s) D_rs for every (p, q) at once
    for (size_t e = tid; e < (size_t)na * na * nc; e += kT) {
      const int pq = (int)(e / nc), f = (int)(e % nc);
      const int p = pq / na, q = pq % na;
      double2 t = zero2();
      for (int r = 0; r < na; ++r)
        for (int s2 = 0; s2 < na; ++s2)
          t = zadd(t, zmul(zscal(I2[p + na * (q + na * (r + na * s2))], 0.5), D[(size_t)(r * na + s2) * nc + f]));
      T[e] = t;
    }
    __syncthreads();
    // S = sum_pq h'_pq E_pq C + sum_pq E_pq T_pq, one determinant per thread
    for (int f = tid; f < nc; f += kT) {
      double2 acc = zero2();
      for (int p = 0; p < na; ++p)
        for (int q = 0; q < na; ++q) acc = zadd(acc, zmul(s_hp[p + na * q], D[(size_t)(p * na + q) * nc + f]));
      for (int p = 0; p < na; ++p)
        for (int q = 0; q < na; ++q) acc = zadd(acc, gather_E(a, v, p, q, T + (size_t)(p * na + q) * nc, f));
      S[f] = acc;
    }
    const double2 fac = a.icomp == 1 ? make_double2(0.0, -a.dtime) : make_double2(-a.dtime, 0.0);
    const double eact = s_eact;
    for (int f = tid; f < nc; f += kT) {
      int d = a.det_of[f];
      if (d < 0) continue;
      a.dcic[d] = zmul(zsub(S[f], zscal(C[f], eact)), fac);
    }
  } else {
    for (int d = tid; d < a.lcic; d += kT) a.dcic[d] = zero2();
  }
  if (a.kout != nullptr) {
    __syncthreads();
    for (int d = tid; d < a.lcic; d += kT) {
      const double2 k = a.dcic[d];
      const bool pre = d == tid && rk_pre;  // inputs already in registers
      // w0 may alias acc: both read before acc is written
      const double2 acc0 = a.acc != nullptr ? (pre ? rk_acc : a.acc[d]) : zero2();
      const double2 w0v = a.next != nullptr ? (pre ? rk_w0 : a.w0[d]) : zero2();
      a.kout[d] = k;
      if (a.acc != nullptr) {
        if (a.w0out != nullptr) a.w0out[d] = acc0;
        a.acc[d] = zadd(acc0, zscal(k, a.dtw));
      }
      if (a.next != nullptr) {
        double2 t = w0v;
        for (int j = 0; j < a.nprev; ++j) t = zadd(t, zscal(pre ? rk_kp[j] : a.kprev[j][d], a.cprev[j]));
        a.next[d] = zadd(t, zscal(k, a.cnew));
      }
    }
  }
}

__global__ void __launch_bounds__(kT) k_island_b(const IslandArgs a) {
  island_b_body(a, hw_blk());
}

void island_lds_plan(IslandArgs& a) {
  a.ci_lds = 0;
  a.i2_lds = 0;
  a.ci_lds_off = 0;
  a.lds_bytes = 0;
  if (a.nact == 0) return;
  const size_t ci = ci_lds_bytes(a.nact, a.nstr_a, a.nstr_b);
  if (ci > 0) {
    a.ci_lds = 1;
    a.ci_lds_off = (ci + 15) / 16 * 16;
  }
  const size_t i2 = (size_t)a.nact * a.nact * a.nact * a.nact * sizeof(double2);
  if (a.ci_lds_off + i2 <= kIslandLds) a.i2_lds = 1;
  a.lds_bytes = a.ci_lds_off + (a.i2_lds ? i2 : 0);
}

void launch_island_a(const IslandArgs& a, hipStream_t s) {
  if (active_recorder()) return record_unsupported("the CI density stage (active orbitals)");
  hipLaunchKernelGGL(k_island_a, dim3(1), dim3(kT), a.ci_lds ? a.ci_lds_off : 0, s, a);
}
void launch_island_b(const IslandArgs& a, hipStream_t s) {
  if (active_recorder()) {
    const int var[4] = {0, 0, 0, 0};
    return record_launch(kRecIslandB, var, dim3(1), kT, a.lds_bytes, a);
  }
  hipLaunchKernelGGL(k_island_b, dim3(1), dim3(kT), a.lds_bytes, s, a);
}

__global__ void __launch_bounds__(kT) k_island_b_batch(const IslandArgs* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(IslandArgs)];
  int have = -1;
  batch_walk(map, NoSkip(), [&](int m, const Blk& bk) {
    const IslandArgs& r = stage_rec(recs, m, raw, have);
    island_b_body(r, bk);
  });
}

bool launch_batch_island(const LaunchRec& p, const void* recs, const BatchMap& map, int grid, size_t lds,
                         hipStream_t s, bool dry) {
  if (p.kind != kRecIslandB || p.block != kT) return false;
  if (!dry)
    hipLaunchKernelGGL(k_island_b_batch, dim3(grid), dim3(kT), lds, s, static_cast<const IslandArgs*>(recs), map);
  return true;
}

}  // namespace td1c
