// Set-up of the FEDVR x spherical-harmonic discretisation for the gfx950
// hprod.  Each table reproduces the one the reference builds (file:line
// given per function) in the layout the kernels read.
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

namespace td1c {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Roots of P'_n on (-1, 1) = interior Gauss-Lobatto nodes (the reference
// takes them as Jacobi-matrix eigenvalues, src/dvr/cldvr.cpp:82-138).
// Newton on (1-x^2) P'_n(x) = n (P_{n-1} - x P_n), deflated by the roots
// already found, long double.
void lobatto_interior(int n, std::vector<long double>& r) {
  r.assign(n > 1 ? n - 1 : 0, 0.0L);
  for (int k = 1; k < n; ++k) {
    long double x = -std::cos(kPi * (long double)k / n);
    for (int it = 0; it < 200; ++it) {
      long double pm = 1.0L, p = x;  // P_{j-1}, P_j
      for (int j = 2; j <= n; ++j) {
        long double pn = ((2 * j - 1) * x * p - (j - 1) * pm) / j;
        pm = p;
        p = pn;
      }
      // f = P'_n * (1 - x^2) / n = P_{n-1} - x P_n ; f' = -n(n+1)/n ... use
      // derivative identity: d/dx[(1-x^2)P'_n] = -n(n+1) P_n
      long double f = pm - x * p;
      long double df = -(long double)(n + 1) * p;
      long double dx = f / df;
      x -= dx;
      if (std::fabs((double)dx) < 1e-19) break;
    }
    r[k - 1] = x;
  }
}

// Normalised associated Legendre functions times sqrt(2 pi) (the SHTOOLS
// norm=4, csphase=-1, cnorm=1 functions used by sph_gen, src/sph/sph_gen.F90:37-79):
// out[l] = sqrt((2l+1)/2 (l-m)!/(l+m)!) P_l^m(x), Condon-Shortley phase included.
void plm_column(int lmax, int m, double x, std::vector<double>& out) {
  out.assign(lmax + 1, 0.0);
  if (m > lmax) return;
  double s = std::sqrt(std::max(0.0, 1.0 - x * x));
  double pmm = std::sqrt(0.5);
  for (int k = 1; k <= m; ++k) pmm *= -s * std::sqrt((2.0 * k + 1.0) / (2.0 * k));
  out[m] = pmm;
  if (m == lmax) return;
  double p1 = x * std::sqrt(2.0 * m + 3.0) * pmm;
  out[m + 1] = p1;
  double p0 = pmm;
  for (int l = m + 2; l <= lmax; ++l) {
    double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
    double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
    double p = a * (x * p1 - b * p0);
    out[l] = p;
    p0 = p1;
    p1 = p;
  }
}

// Gauss-Legendre nodes (ascending) and weights.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double z = std::cos(kPi * (i + 0.75L) / (n + 0.5L));
    long double dp = 1.0L;
    for (int it = 0; it < 200; ++it) {
      long double pm = 1.0L, p = z;
      for (int j = 2; j <= n; ++j) {
        long double pn = ((2 * j - 1) * z * p - (j - 1) * pm) / j;
        pm = p;
        p = pn;
      }
      if (n == 1) { p = z; pm = 1.0L; }
      dp = n * (z * p - pm) / (z * z - 1.0L);
      long double dz = p / dp;
      z -= dz;
      if (std::fabs((double)dz) < 1e-19) {
        // one more derivative at the converged node
        pm = 1.0L; p = z;
        for (int j = 2; j <= n; ++j) {
          long double pn = ((2 * j - 1) * z * p - (j - 1) * pm) / j;
          pm = p;
          p = pn;
        }
        dp = n * (z * p - pm) / (z * z - 1.0L);
        break;
      }
    }
    x[i] = (double)(-z);
    x[n - 1 - i] = (double)z;
    double wi = (double)(2.0L / ((1.0L - z * z) * dp * dp));
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
}

long double legendre_value(int n, long double x) {
  if (n == 0) return 1.0L;
  long double pm = 1.0L, p = x;
  for (int j = 2; j <= n; ++j) {
    long double pn = ((2 * j - 1) * x * p - (j - 1) * pm) / j;
    pm = p;
    p = pn;
  }
  return p;
}

// Dense inverse of a small SPD matrix (Gauss-Jordan with partial pivoting).
bool invert(int m, std::vector<double>& a) {
  std::vector<double> inv(m * m, 0.0);
  for (int i = 0; i < m; ++i) inv[i * m + i] = 1.0;
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(a[r * m + c]) > std::fabs(a[piv * m + c])) piv = r;
    if (a[piv * m + c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < m; ++k) {
        std::swap(a[piv * m + k], a[c * m + k]);
        std::swap(inv[piv * m + k], inv[c * m + k]);
      }
    double d = 1.0 / a[c * m + c];
    for (int k = 0; k < m; ++k) {
      a[c * m + k] *= d;
      inv[c * m + k] *= d;
    }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      double f = a[r * m + c];
      if (f == 0.0) continue;
      for (int k = 0; k < m; ++k) {
        a[r * m + k] -= f * a[c * m + k];
        inv[r * m + k] -= f * inv[c * m + k];
      }
    }
  }
  a.swap(inv);
  return true;
}

// Dense inverse of a small complex matrix (Gauss-Jordan with partial
// pivoting on |re| + |im|).
bool invert_c(int m, std::vector<std::complex<double>>& a) {
  using cd = std::complex<double>;
  std::vector<cd> inv(m * m, cd(0.0, 0.0));
  for (int i = 0; i < m; ++i) inv[i * m + i] = 1.0;
  auto mag = [](cd z) { return std::fabs(z.real()) + std::fabs(z.imag()); };
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (mag(a[r * m + c]) > mag(a[piv * m + c])) piv = r;
    if (a[piv * m + c] == cd(0.0, 0.0)) return false;
    if (piv != c)
      for (int k = 0; k < m; ++k) {
        std::swap(a[piv * m + k], a[c * m + k]);
        std::swap(inv[piv * m + k], inv[c * m + k]);
      }
    const cd d = 1.0 / a[c * m + c];
    for (int k = 0; k < m; ++k) {
      a[c * m + k] *= d;
      inv[c * m + k] *= d;
    }
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const cd f = a[r * m + c];
      if (f == cd(0.0, 0.0)) continue;
      for (int k = 0; k < m; ++k) {
        a[r * m + k] -= f * a[c * m + k];
        inv[r * m + k] -= f * inv[c * m + k];
      }
    }
  }
  a.swap(inv);
  return true;
}

int popc(uint32_t v) { return __builtin_popcount(v); }

}  // namespace

// Element nodes / weights (cldvr::gen, src/dvr/cldvr.cpp:57-221): Gauss-Lobatto
// points of order n mapped onto [edges[e], edges[e+1]]; global points 0..n*nfe
// with the weights of the two elements summed at each bridge
// (clfedvr::gen_grid, src/fedvr/clfedvr.cpp:400-511).
void fedvr_nodes(int nfe, int n, const double* edges, std::vector<double>& ex, std::vector<double>& ew,
                 std::vector<double>& xrad, std::vector<double>& wrad) {
  const int np = n + 1, nrad = n * nfe;
  std::vector<long double> inner;
  lobatto_interior(n, inner);
  std::vector<double> ref_x(np), ref_w(np);
  ref_x[0] = -1.0; ref_x[n] = 1.0;
  for (int k = 1; k < n; ++k) ref_x[k] = (double)inner[k - 1];
  const double scale = 2.0 / ((double)n * (n + 1.0));
  ref_w[0] = ref_w[n] = scale;
  for (int k = 1; k < n; ++k) {
    double pv = (double)legendre_value(n, std::fabs(ref_x[k]));
    ref_w[k] = scale / (pv * pv);
  }
  ex.assign((size_t)nfe * np, 0); ew.assign((size_t)nfe * np, 0);
  for (int e = 0; e < nfe; ++e) {
    double half = 0.5 * (edges[e + 1] - edges[e]), mid = 0.5 * (edges[e + 1] + edges[e]);
    for (int k = 0; k < np; ++k) {
      ex[(size_t)e * np + k] = ref_x[k] * half + mid;
      ew[(size_t)e * np + k] = ref_w[k] * half;
    }
  }
  xrad.assign(nrad + 1, 0); wrad.assign(nrad + 1, 0);
  for (int ib = 0; ib <= nrad; ++ib) {
    int e = std::min(ib / n, nfe - 1), m = ib - e * n;
    xrad[ib] = ex[(size_t)e * np + m];
    wrad[ib] = ew[(size_t)e * np + m];
  }
  for (int e = 1; e < nfe; ++e) wrad[e * n] += ew[(size_t)(e - 1) * np + n];
}

bool Setup::build(const td1c_problem& p) {
  if (p.nfe < 2 || p.ndvr < 2 || !p.edges) { error = "need nfe >= 2, ndvr >= 2 and element edges"; return false; }
  if (p.lmax1 < 0 || p.lmax2 < 0) { error = "bad lmax1/lmax2"; return false; }
  if (p.ioorot != 0) { error = "only oorot_type = 0 is provided"; return false; }
  if (p.isplit != 0 && p.isplit != 1) { error = "split_type must be 0 (full hprod) or 1 (clh1itr one-body half steps)"; return false; }
  if (p.isplit == 1 && p.iprojfc == 1 && p.nfcore2 + p.nfcore1 > 0) {
    error = "split_type = 1 with frozen cores needs iprojfc = 2 (h1itr_prop's frozen-core projection is not provided)";
    return false;
  }
  if (p.iprojfc != 1 && p.iprojfc != 2) { error = "iprojfc must be 1 or 2"; return false; }
  if (!p.jfc_implicit || !p.xfc_implicit) { error = "only jfc_implicit = xfc_implicit = true are provided"; return false; }
  nfe = p.nfe; n = p.ndvr; nrad = nfe * n; nr = nrad - 1;
  lmax1 = p.lmax1; lmax2 = p.lmax2; L1 = lmax1 + 1; L2 = lmax2 + 1;
  nlat = p.nlat > 0 ? p.nlat : lmax2 + 1;
  znuc = p.znuc;
  nfcore2 = p.nfcore2; nfcore1 = p.nfcore1; nfcore = nfcore2 + nfcore1;
  ndcore = p.ndcore; ncore = nfcore + ndcore; nact = p.nact; nfun = ncore + nact; ndyn = nfun - nfcore;
  if (nfun <= 0 || ndyn <= 0) { error = "need at least one dynamical orbital"; return false; }
  if (nact > 16) { error = "nact > 16 is not provided"; return false; }
  mval.assign(nfun, 0); lval.assign(nfun, 0);
  int mmax = 0;
  for (int i = 0; i < nfun; ++i) {
    mval[i] = p.mval ? p.mval[i] : 0;
    lval[i] = p.lval ? p.lval[i] : 0;
    mmax = std::max(mmax, std::abs(mval[i]));
  }
  mmax1 = p.mmax1 >= 0 ? p.mmax1 : mmax;
  mmax2 = p.mmax2 >= 0 ? p.mmax2 : 2 * mmax1;
  if (mmax > mmax1 || 2 * mmax > mmax2) { error = "mmax1/mmax2 smaller than the orbital m values need"; return false; }
  nel[0] = p.nel_alpha; nel[1] = p.nel_beta; nel[2] = nel[0] + nel[1];
  nelact[0] = nel[0] - ncore; nelact[1] = nel[1] - ncore; nelact[2] = nelact[0] + nelact[1];
  if (nelact[0] < 0 || nelact[1] < 0 || nelact[0] > nact || nelact[1] > nact) { error = "electron count does not fit the orbital classes"; return false; }
  mtot = p.mtot;
  icomp = p.icomp; igauge = p.igauge; isplit = p.isplit; ioorot = p.ioorot; iprojfc = p.iprojfc;
  jfc_implicit = p.jfc_implicit; xfc_implicit = p.xfc_implicit;
  throcc1 = p.throcc1; throcc2 = p.throcc2;
  precision = p.precision;
  if (precision != 0 && precision != 1) { error = "precision must be 0 (complex128) or 1 (complex64 mean field)"; return false; }
  lconst = p.lconst;
  if (lconst < 0 || lconst > 2) { error = "lconst must be 0, 1 (all orbitals) or 2 (cores)"; return false; }
  x0.resize(nfe); x1.resize(nfe);
  for (int e = 0; e < nfe; ++e) { x0[e] = p.edges[e]; x1[e] = p.edges[e + 1]; }
  rmask = p.rmask > 0 ? p.rmask : 0.8 * x1[nfe - 1];
  mask_order = p.mask_cos_order > 0 ? p.mask_cos_order : 4.0;
  nradfc = p.nradfc > 0 ? p.nradfc : nr;
  if (nradfc != nr && nfcore1 > 0) { error = "frozen-core exchange pairs need nradfc = nrad-1"; return false; }

  // ---- element DVR and global points ----
  const int np = n + 1;
  fedvr_nodes(nfe, n, p.edges, ex, ew, xrad, wrad);
  eds.assign((size_t)nfe * np * np, 0);
  for (int e = 0; e < nfe; ++e) {
    const double* xe = &ex[(size_t)e * np];
    const double* we = &ew[(size_t)e * np];
    double* ds = &eds[(size_t)e * np * np];  // ds[np*m + k] = f_m'(x_k)
    for (int m = 0; m < np; ++m)
      for (int k = 0; k < np; ++k) {
        if (m == k) { ds[np * m + k] = 0.0; continue; }
        double v = 1.0 / (xe[m] - xe[k]);
        for (int l = 0; l < np; ++l)
          if (l != m && l != k) v *= (xe[k] - xe[l]) / (xe[m] - xe[l]);
        ds[np * m + k] = v;
      }
    ds[0] = -0.5 / we[0];
    ds[np * np - 1] = 0.5 / we[n];
  }
  mask.assign(nrad + 1, 0);
  // mask (clfedvr::gen_mask, :361-398)
  {
    double order = 1.0 / mask_order, rul = x1[nfe - 1];
    for (int ib = 1; ib < nrad; ++ib) {
      double a = std::fabs(xrad[ib]);
      mask[ib] = a <= rmask ? 1.0 : std::pow(std::cos(0.5 * kPi * (a - rmask) / (rul - rmask)), order);
    }
  }
  // kinetic and derivative bands (clfedvr::gen_radk :676-757, gen_radp :759-792)
  const size_t bw = 2 * (size_t)n + 1;
  radk.assign(bw * (nrad + 1), 0); radk0.assign(bw * (nrad + 1), 0); radp.assign(bw * (nrad + 1), 0);
  for (int e = 0; e < nfe; ++e) {
    const double* we = &ew[(size_t)e * np];
    const double* ds = &eds[(size_t)e * np * np];
    for (int m = 0; m < np; ++m)
      for (int k = 0; k < np; ++k) {
        if (e == 0 && (m == 0 || k == 0)) continue;
        if (e == nfe - 1 && (m == n || k == n)) continue;
        int mb = e * n + m, kb = e * n + k;
        size_t at = bw * mb + (n + kb - mb);
        double t = 0, t0 = 0;
        for (int l = 0; l < np; ++l) {
          double c = we[l] * ds[np * m + l] * ds[np * k + l];
          t += c;
          if (!(e == 0 && l == 0)) t0 += c;
        }
        double nrm = std::sqrt(wrad[mb] * wrad[kb]);
        radk[at] += 0.5 * t / nrm;
        radk0[at] += 0.5 * t0 / nrm;
        radp[at] += we[m] * ds[np * k + m] / nrm;
      }
  }

  // element-local kinetic / derivative matrices (the per-element terms that
  // gen_radk / gen_radp sum into the global bands)
  eK0.assign((size_t)nfe * np * np, 0); eK.assign((size_t)nfe * np * np, 0); eP.assign((size_t)nfe * np * np, 0);
  for (int e = 0; e < nfe; ++e) {
    const double* we = &ew[(size_t)e * np];
    const double* ds = &eds[(size_t)e * np * np];
    for (int m = 0; m < np; ++m)
      for (int k = 0; k < np; ++k) {
        if (e == 0 && (m == 0 || k == 0)) continue;
        if (e == nfe - 1 && (m == n || k == n)) continue;
        int mb = e * n + m, kb = e * n + k;
        double t = 0, t0 = 0;
        for (int l = 0; l < np; ++l) {
          double c = we[l] * ds[np * m + l] * ds[np * k + l];
          t += c;
          if (!(e == 0 && l == 0)) t0 += c;
        }
        double nrm = std::sqrt(wrad[mb] * wrad[kb]);
        size_t at = ((size_t)e * np + k) * np + m;
        eK0[at] = 0.5 * t / nrm;
        eK[at] = 0.5 * t0 / nrm;
        eP[at] = we[m] * ds[np * k + m] / nrm;
      }
  }
  {
    // keep each distinct (eK0, eK, eP) element block once
    const size_t bs = (size_t)np * np;
    std::map<std::string, int> seen;
    std::vector<double> u0, u1, u2;
    emap.assign(nfe, 0);
    for (int e = 0; e < nfe; ++e) {
      std::string key(reinterpret_cast<const char*>(&eK0[e * bs]), bs * sizeof(double));
      key.append(reinterpret_cast<const char*>(&eK[e * bs]), bs * sizeof(double));
      key.append(reinterpret_cast<const char*>(&eP[e * bs]), bs * sizeof(double));
      auto it = seen.find(key);
      if (it != seen.end()) {
        emap[e] = it->second;
        continue;
      }
      const int u = (int)seen.size();
      seen.emplace(std::move(key), u);
      emap[e] = u;
      u0.insert(u0.end(), eK0.begin() + e * bs, eK0.begin() + (e + 1) * bs);
      u1.insert(u1.end(), eK.begin() + e * bs, eK.begin() + (e + 1) * bs);
      u2.insert(u2.end(), eP.begin() + e * bs, eP.begin() + (e + 1) * bs);
    }
    eK0.swap(u0);
    eK.swap(u1);
    eP.swap(u2);
  }

  // ---- kernel-side one-body tables ----
  xr.assign(nr, 0);
  for (int i = 0; i < nr; ++i) xr[i] = xrad[i + 1];
  bandT_k.assign(bw * nr, 0); bandT_k0.assign(bw * nr, 0); bandT_p.assign(bw * nr, 0);
  for (int i = 0; i < nr; ++i)
    for (int d = 0; d < (int)bw; ++d) {
      int j = i + d - n;
      if (j < 0 || j >= nr) continue;
      bandT_k[(size_t)d * nr + i] = radk[bw * (i + 1) + d];
      bandT_k0[(size_t)d * nr + i] = radk0[bw * (i + 1) + d];
      bandT_p[(size_t)d * nr + i] = radp[bw * (i + 1) + d];
    }
  // diagonal of tmat (bas_gen_tmat, src/bas/bas_gen_tmat.F90:63-118): the
  // Coulomb+centrifugal value first, then the kinetic diagonal added
  tdiag.assign((size_t)2 * L1 * nr, 0);
  tzor.assign((size_t)2 * L1 * nr, 0);
  for (int l = 0; l < L1; ++l)
    for (int i = 0; i < nr; ++i) {
      double oor = 1.0 / xr[i];
      double zor = oor * (-znuc + (double)(l * (l + 1)) * oor * 0.5);
      double kd = (l == 0 ? radk : radk0)[bw * (i + 1) + n];
      tdiag[2 * ((size_t)l * nr + i)] = zor + kd;
      tzor[2 * ((size_t)l * nr + i)] = zor;
    }
  // alph_lm (bas_gen_alph, src/bas/bas_gen_alph.F90:14-25)
  alph.assign((size_t)(2 * mmax1 + 1) * L1, 0);
  for (int m = -mmax1; m <= mmax1; ++m)
    for (int l = std::abs(m); l <= lmax1; ++l) {
      double dl = l, dm = m;
      alph[(size_t)(m + mmax1) * L1 + l] =
          std::sqrt(((dl + 1) * (dl + 1) - dm * dm) / ((2 * dl + 1) * (2 * dl + 3)));
    }

  // ---- angular quadrature and transform matrices (sph_gen) ----
  gauss_legendre(nlat, cost, wlat);
  const int M1 = 2 * mmax1 + 1, M2 = 2 * mmax2 + 1;
  mat_l2g.assign((size_t)M1 * L1 * nlat, 0);
  mat_g2l.assign((size_t)M1 * nlat * L1, 0);
  mat_g2L.assign((size_t)M2 * nlat * L2, 0);
  mat_L2g.assign((size_t)M2 * L2 * nlat, 0);
  std::vector<double> col;
  int lmx = std::max(lmax1, lmax2);
  for (int t = 0; t < nlat; ++t)
    for (int am = 0; am <= std::max(mmax1, mmax2); ++am) {
      plm_column(lmx, am, cost[t], col);
      for (int sg = -1; sg <= 1; sg += 2) {
        if (am == 0 && sg < 0) continue;
        int m = sg * am;  // negative m: same table, no sign (sph_gen.F90:85-100)
        if (am <= mmax1)
          for (int l = am; l <= lmax1; ++l) {
            mat_l2g[((size_t)(m + mmax1) * L1 + l) * nlat + t] = col[l];
            mat_g2l[((size_t)(m + mmax1) * nlat + t) * L1 + l] = col[l] * wlat[t];
          }
        if (am <= mmax2)
          for (int l = am; l <= lmax2; ++l) {
            mat_L2g[((size_t)(m + mmax2) * L2 + l) * nlat + t] = col[l];
            mat_g2L[((size_t)(m + mmax2) * nlat + t) * L2 + l] = col[l] * wlat[t];
          }
      }
    }

  // ---- Poisson (bas_gen_d2ll / bas_gen_d2fac / hprod_mkv2mf_poissonp) ----
  const double r2pi = 0.5 / kPi;
  d2fac1.assign(L2, 0); d2invr.assign((size_t)L2 * nr, 0); d2rpl0.assign((size_t)L2 * nr, 0);
  d2rpl1.assign((size_t)L2 * nr, 0); d2fac2.assign((size_t)L2 * nr, 0);
  for (int L = 0; L < L2; ++L) {
    d2fac1[L] = 1.0 / std::pow(xrad[nr + 1], 2 * L + 1);  // 0 once R^(2L+1) overflows
    for (int i = 0; i < nr; ++i) {
      double x = xr[i], sw = std::sqrt(wrad[i + 1]);
      size_t k = (size_t)L * nr + i;
      d2rpl0[k] = std::pow(x, L) * r2pi;
      d2rpl1[k] = std::pow(x, L + 1) * sw;
      d2invr[k] = (2 * L + 1) / (x * sw) * r2pi;
      d2fac2[k] = 4.0 * kPi / (2 * L + 1) / (x * sw);
    }
  }
  // The operator is A_L = 2 K + diag(L(L+1)/r^2) on the interior points
  // (src/bas/bas_gen_d2ll.F90:25-57).  Interior unknowns of element e are its
  // local points 1..n-1 (global i = e*n + m - 1); bridge k (1..nfe-1) is
  // global i = k*n - 1.  Eliminating the element interiors leaves a
  // tridiagonal Schur system on the bridges.
  ni = n - 1;
  const int nb = nfe - 1;
  pG.assign((size_t)L2 * nfe * ni * ni, 0);
  pCl.assign((size_t)L2 * nfe * ni, 0);
  pCr.assign((size_t)L2 * nfe * ni, 0);
  pWl.assign((size_t)L2 * nr, 0);
  pWr.assign((size_t)L2 * nr, 0);
  pSd.assign((size_t)L2 * std::max(nb, 1), 0);
  pSl.assign((size_t)L2 * std::max(nb, 1), 0);
  pSu.assign((size_t)L2 * std::max(nb, 1), 0);
  pcr_steps = 0;
  while ((1 << pcr_steps) < nb) ++pcr_steps;
  pcrA.assign((size_t)L2 * std::max(pcr_steps, 1) * std::max(nb, 1), 0);
  pcrG.assign((size_t)L2 * std::max(pcr_steps, 1) * std::max(nb, 1), 0);
  pcrBinv.assign((size_t)L2 * std::max(nb, 1), 0);
  auto A = [&](int L, int i, int j) -> double {  // interior indices
    if (std::abs(i - j) > n) return 0.0;
    double v = 2.0 * radk[bw * (i + 1) + (n + j - i)];
    if (i == j) v += (double)(L * (L + 1)) / (xr[i] * xr[i]);
    return v;
  };
  std::vector<double> blk(ni * ni), sdiag(nb), soff(nb);
  for (int L = 0; L < L2; ++L) {
    for (int k = 0; k < nb; ++k) {
      int gi = (k + 1) * n - 1;
      sdiag[k] = A(L, gi, gi);
      soff[k] = k + 1 < nb ? A(L, gi, gi + n) : 0.0;
    }
    for (int e = 0; e < nfe; ++e) {
      int base = e * n;  // interior local m=1..n-1 -> global i = base + m - 1
      for (int a = 0; a < ni; ++a)
        for (int b = 0; b < ni; ++b) blk[a * ni + b] = A(L, base + a, base + b);
      if (!invert(ni, blk)) { error = "singular element block in the Poisson operator"; return false; }
      double* G = &pG[(((size_t)L * nfe) + e) * ni * ni];
      std::copy(blk.begin(), blk.end(), G);
      int bl = e >= 1 ? e * n - 1 : -1;             // left bridge global index
      int br = e <= nfe - 2 ? (e + 1) * n - 1 : -1; // right bridge global index
      double* cl = &pCl[(((size_t)L * nfe) + e) * ni];
      double* cr = &pCr[(((size_t)L * nfe) + e) * ni];
      for (int a = 0; a < ni; ++a) {
        cl[a] = bl >= 0 ? A(L, bl, base + a) : 0.0;
        cr[a] = br >= 0 ? A(L, br, base + a) : 0.0;
      }
      // W = G * A(I, bridge); symmetric A: A(I, b) = C^T
      for (int a = 0; a < ni; ++a) {
        double wl = 0, wr = 0;
        for (int b = 0; b < ni; ++b) { wl += G[a * ni + b] * cl[b]; wr += G[a * ni + b] * cr[b]; }
        pWl[(size_t)L * nr + base + a] = wl;
        pWr[(size_t)L * nr + base + a] = wr;
      }
      // Schur updates: S(k,k) -= C G C^T for each adjacent bridge, S(k,k+1) -= Cl G Cr^T
      double gll = 0, grr = 0, glr = 0;
      for (int a = 0; a < ni; ++a) {
        gll += cl[a] * pWl[(size_t)L * nr + base + a];
        grr += cr[a] * pWr[(size_t)L * nr + base + a];
        glr += cl[a] * pWr[(size_t)L * nr + base + a];
      }
      if (bl >= 0) sdiag[e - 1] -= gll;
      if (br >= 0) sdiag[e] -= grr;
      if (bl >= 0 && br >= 0) soff[e - 1] -= glr;
    }
    // parallel cyclic reduction coefficients of S (symmetric: sub = super)
    {
      const int K = pcr_steps;
      std::vector<double> a(nb), b(nb), c(nb), na(nb), nbv(nb), nc(nb);
      for (int i = 0; i < nb; ++i) {
        a[i] = i > 0 ? soff[i - 1] : 0.0;
        b[i] = sdiag[i];
        c[i] = i + 1 < nb ? soff[i] : 0.0;
      }
      for (int k = 0, h = 1; k < K; ++k, h *= 2) {
        for (int i = 0; i < nb; ++i) {
          const double al = i - h >= 0 ? -a[i] / b[i - h] : 0.0;
          const double ga = i + h < nb ? -c[i] / b[i + h] : 0.0;
          pcrA[((size_t)L * K + k) * nb + i] = al;
          pcrG[((size_t)L * K + k) * nb + i] = ga;
          na[i] = i - h >= 0 ? al * a[i - h] : 0.0;
          nc[i] = i + h < nb ? ga * c[i + h] : 0.0;
          nbv[i] = b[i] + (i - h >= 0 ? al * c[i - h] : 0.0) + (i + h < nb ? ga * a[i + h] : 0.0);
        }
        a.swap(na);
        b.swap(nbv);
        c.swap(nc);
      }
      for (int i = 0; i < nb; ++i) {
        if (!(b[i] > 0)) { error = "Poisson Schur system: cyclic reduction lost positivity"; return false; }
        pcrBinv[(size_t)L * nb + i] = 1.0 / b[i];
      }
    }
    // Thomas factorisation of the symmetric tridiagonal S
    double* Sd = &pSd[(size_t)L * nb];
    double* Sl = &pSl[(size_t)L * nb];
    double* Su = &pSu[(size_t)L * nb];
    for (int k = 0; k < nb; ++k) {
      double piv = sdiag[k];
      if (k > 0) {
        Sl[k] = soff[k - 1] / Sd[k - 1];
        piv -= Sl[k] * soff[k - 1];
      }
      if (!(piv > 0)) { error = "Poisson Schur system is not positive definite"; return false; }
      Sd[k] = piv;
      Su[k] = soff[k];
    }
  }

  // ---- orbital pairs (hprod_mkrho2p 'dyn' then 'fc1dyn') ----
  pairs.clear();
  for (int i = nfcore; i < nfun; ++i)
    for (int j = i; j < nfun; ++j) pairs.push_back({j, i, 0});
  for (int i = nfcore2; i < nfcore; ++i)
    for (int j = nfcore; j < nfun; ++j) pairs.push_back({j, i, 1});
  pair_of.assign((size_t)nfun * nfun, -1);
  pair_conj.assign((size_t)nfun * nfun, 0);
  for (size_t q = 0; q < pairs.size(); ++q) {
    int j = pairs[q].j, i = pairs[q].i;
    pair_of[(size_t)j * nfun + i] = (int)q;
    pair_of[(size_t)i * nfun + j] = (int)q;
    if (i != j) pair_conj[(size_t)i * nfun + j] = 1;  // v(i,j) = conj(v(j,i))
  }

  return build_ci();
}

// CAS determinant space in the reference's layout (ormas_init with one
// subspace, src/ormas/ormas_init.F90:32-53).  Strings of one spin are first
// numbered by arc weights (ormas_arcwgt, src/ormas/ormas_arcwgt.F90:100-104;
// ormas_occvec, src/ormas/ormas_occvec.F90:17-25): the string with occupied
// orbitals o_1 < ... < o_k has rank sum_k C(o_k - 1, k), which is the
// ascending order of the occupation bit patterns (bit o = orbital o).
// ormas_madapt_str (src/ormas/ormas_madapt.F90:170-246) then reorders them by
// their total M (cores included), keeping that order within one M, and the CI
// vector holds the determinants with M_alpha + M_beta = mtot, beta string
// major and alpha fastest: index ntot_alph_beta(beta) + alpha (:381-435).
bool Setup::build_ci() {
  if (nact == 0) {
    nstr_a = nstr_b = 1;
    lcic = 1;
    cic0 = 0;
    return true;
  }
  int mcore = 0;
  for (int k = 0; k < ncore; ++k) mcore += mval[k];
  auto mstr = [&](uint32_t v) {
    int m = mcore;
    for (int k = 0; k < nact; ++k)
      if (v >> k & 1u) m += mval[ncore + k];
    return m;
  };
  // returns the new position of the first (rank-0) string
  auto strings = [&](int ne, std::vector<uint32_t>& s) {
    s.clear();
    for (uint32_t v = 0; v < (1u << nact); ++v)
      if (popc(v) == ne) s.push_back(v);
    const uint32_t first = s.front();
    std::stable_sort(s.begin(), s.end(), [&](uint32_t a, uint32_t b) { return mstr(a) < mstr(b); });
    return (int)(std::find(s.begin(), s.end(), first) - s.begin());
  };
  const int first_a = strings(nelact[0], str_a);
  const int first_b = strings(nelact[1], str_b);
  nstr_a = (int)str_a.size();
  nstr_b = (int)str_b.size();
  det_of.assign((size_t)nstr_a * nstr_b, -1);
  std::vector<int> ntot(nstr_b);
  lcic = 0;
  for (int ib = 0; ib < nstr_b; ++ib) {
    // the alpha strings allowed with beta string ib are one contiguous range
    int ll = -1, cnt = 0;
    for (int ia = 0; ia < nstr_a; ++ia)
      if (mstr(str_a[ia]) + mstr(str_b[ib]) == mtot) {
        if (ll < 0) ll = ia;
        ++cnt;
      }
    ntot[ib] = lcic - ll;  // ll = -1 (llstr_alph_beta = 0) when none is allowed
    for (int ia = ll; ia >= 0 && ia < ll + cnt; ++ia) det_of[(size_t)ib * nstr_a + ia] = ntot[ib] + ia;
    lcic += cnt;
  }
  if (lcic == 0) { error = "no determinant has M = mtot"; return false; }
  // ormas_cic0 (src/ormas/ormas_cic0.F90:24): the determinant of the first
  // alpha and beta strings, ntot_alph_beta(map1to2_beta(1)) + map1to2_alph(1)
  cic0 = ntot[first_b] + first_a;
  auto table = [&](const std::vector<uint32_t>& s, std::vector<int>& ex, std::vector<int>& sg) {
    const int ns = (int)s.size();
    std::vector<int> pos(1u << nact, -1);
    for (int a = 0; a < ns; ++a) pos[s[a]] = a;
    ex.assign((size_t)ns * nact * nact, -1);
    sg.assign((size_t)ns * nact * nact, 0);
    for (int a = 0; a < ns; ++a)
      for (int pp = 0; pp < nact; ++pp)
        for (int q = 0; q < nact; ++q) {
          const uint32_t v = s[a];
          if (!(v >> q & 1u)) continue;
          uint32_t t;
          int sign = 1;
          if (pp == q) {
            t = v;
          } else {
            if (v >> pp & 1u) continue;
            // ormas_sgn1x (src/ormas/ormas_sgn1x.F90): (-1)^(occupied orbitals strictly between)
            const int lo = std::min(pp, q), hi = std::max(pp, q);
            const uint32_t between = v & (((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u));
            sign = (popc(between) & 1) ? -1 : 1;
            t = (v & ~(1u << q)) | (1u << pp);
          }
          ex[((size_t)a * nact + pp) * nact + q] = pos[t];
          sg[((size_t)a * nact + pp) * nact + q] = sign;
        }
  };
  table(str_a, exA, sgA);
  table(str_b, exB, sgB);
  return true;
}

// Crank-Nicolson factors of clh1itr (h1itr_gen_cninv,
// src/prop/h1itr_gen_cninv.F90:2-58): A_l = 1 + fac T_l on the interior
// points, with the element / bridge split of the Poisson operator.  The
// reference factors the whole band by zgbtrf; the solution is the same.
bool build_cn(const Setup& S, const double* v2jfc, double fac_re, double fac_im, CnTables& T, std::string& error) {
  using cd = std::complex<double>;
  const int n = S.n, nr = S.nr, nfe = S.nfe, L1 = S.L1, ni = n - 1, nb = nfe - 1;
  const size_t bw = 2 * (size_t)n + 1;
  const cd fac(fac_re, fac_im);
  T.L1 = L1; T.nfe = nfe; T.ni = ni; T.nr = nr; T.nb = nb;
  T.K = 0;
  while ((1 << T.K) < nb) ++T.K;
  const int K = T.K;
  T.G.assign((size_t)2 * L1 * nfe * ni * ni, 0.0);
  T.Wl.assign((size_t)2 * L1 * nr, 0.0);
  T.Wr.assign((size_t)2 * L1 * nr, 0.0);
  T.pa.assign((size_t)2 * L1 * std::max(K, 1) * std::max(nb, 1), 0.0);
  T.pg.assign((size_t)2 * L1 * std::max(K, 1) * std::max(nb, 1), 0.0);
  T.binv.assign((size_t)2 * L1 * std::max(nb, 1), 0.0);
  auto put = [](std::vector<double>& v, size_t q, cd z) { v[2 * q] = z.real(); v[2 * q + 1] = z.imag(); };
  for (int l = 0; l < L1; ++l) {
    const std::vector<double>& kb = l == 0 ? S.radk : S.radk0;
    auto A = [&](int i, int j) -> cd {  // interior indices
      if (std::abs(i - j) > n) return cd(0.0, 0.0);
      cd t;
      if (i == j) {
        const size_t q = (size_t)l * nr + i;
        t = cd(S.tdiag[2 * q], S.tdiag[2 * q + 1]);
        if (v2jfc) t += cd(v2jfc[2 * i], v2jfc[2 * i + 1]);
      } else {
        t = kb[bw * (i + 1) + (n + j - i)];
      }
      return (i == j ? cd(1.0, 0.0) : cd(0.0, 0.0)) + fac * t;
    };
    std::vector<cd> sdiag(std::max(nb, 1)), soff(std::max(nb, 1)), blk((size_t)ni * ni), wl(ni), wr(ni);
    for (int k = 0; k < nb; ++k) {
      const int gi = (k + 1) * n - 1;
      sdiag[k] = A(gi, gi);
      soff[k] = k + 1 < nb ? A(gi, gi + n) : cd(0.0, 0.0);
    }
    for (int e = 0; e < nfe; ++e) {
      const int base = e * n;
      for (int a = 0; a < ni; ++a)
        for (int b = 0; b < ni; ++b) blk[(size_t)a * ni + b] = A(base + a, base + b);
      if (!invert_c(ni, blk)) {
        error = "singular element block of the Crank-Nicolson operator";
        return false;
      }
      // stored transposed, [k][i] = G(i, k), so that lane i reads a row of G with stride-1 loads
      for (int a = 0; a < ni; ++a)
        for (int k = 0; k < ni; ++k) put(T.G, ((size_t)l * nfe + e) * ni * ni + (size_t)k * ni + a, blk[(size_t)a * ni + k]);
      const int bl = e >= 1 ? e * n - 1 : -1, br = e <= nfe - 2 ? (e + 1) * n - 1 : -1;
      std::vector<cd> cl(ni), cr(ni);
      for (int a = 0; a < ni; ++a) {
        cl[a] = bl >= 0 ? A(bl, base + a) : cd(0.0, 0.0);
        cr[a] = br >= 0 ? A(br, base + a) : cd(0.0, 0.0);
      }
      for (int a = 0; a < ni; ++a) {
        cd sl(0.0, 0.0), sr(0.0, 0.0);
        for (int b = 0; b < ni; ++b) {
          sl += blk[(size_t)a * ni + b] * cl[b];
          sr += blk[(size_t)a * ni + b] * cr[b];
        }
        wl[a] = sl;
        wr[a] = sr;
        put(T.Wl, (size_t)l * nr + base + a, sl);
        put(T.Wr, (size_t)l * nr + base + a, sr);
      }
      cd gll(0.0, 0.0), grr(0.0, 0.0), glr(0.0, 0.0);
      for (int a = 0; a < ni; ++a) {
        gll += cl[a] * wl[a];
        grr += cr[a] * wr[a];
        glr += cl[a] * wr[a];
      }
      if (bl >= 0) sdiag[e - 1] -= gll;
      if (br >= 0) sdiag[e] -= grr;
      if (bl >= 0 && br >= 0) soff[e - 1] -= glr;
    }
    // parallel cyclic reduction of the symmetric tridiagonal bridge system
    std::vector<cd> a(std::max(nb, 1)), b(std::max(nb, 1)), c(std::max(nb, 1)), na(a), nbv(b), nc(c);
    for (int i = 0; i < nb; ++i) {
      a[i] = i > 0 ? soff[i - 1] : cd(0.0, 0.0);
      b[i] = sdiag[i];
      c[i] = i + 1 < nb ? soff[i] : cd(0.0, 0.0);
    }
    for (int k = 0, h = 1; k < K; ++k, h *= 2) {
      for (int i = 0; i < nb; ++i) {
        const cd al = i - h >= 0 ? -a[i] / b[i - h] : cd(0.0, 0.0);
        const cd ga = i + h < nb ? -c[i] / b[i + h] : cd(0.0, 0.0);
        put(T.pa, ((size_t)l * K + k) * nb + i, al);
        put(T.pg, ((size_t)l * K + k) * nb + i, ga);
        na[i] = i - h >= 0 ? al * a[i - h] : cd(0.0, 0.0);
        nc[i] = i + h < nb ? ga * c[i + h] : cd(0.0, 0.0);
        nbv[i] = b[i] + (i - h >= 0 ? al * c[i - h] : cd(0.0, 0.0)) + (i + h < nb ? ga * a[i + h] : cd(0.0, 0.0));
      }
      a.swap(na);
      b.swap(nbv);
      c.swap(nc);
    }
    for (int i = 0; i < nb; ++i) {
      if (!(std::abs(b[i]) > 1e-300)) {
        error = "Crank-Nicolson bridge system: cyclic reduction met a zero pivot";
        return false;
      }
      put(T.binv, (size_t)l * nb + i, 1.0 / b[i]);
    }
  }
  return true;
}

}  // namespace td1c
