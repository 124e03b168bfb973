/*
 * td1c_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 CPU restatement of td1c's orbital Hamiltonian product
 * (hprod_htot, src/hprod/hprod_htot.F90:2-208), its set-up (FEDVR grid,
 * Gauss-Legendre / associated-Legendre tables, operator matrices) and the
 * explicit Runge-Kutta step clcrk::prop (src/prop/clcrk.cpp:55-87).
 * It is the checker the GPU library is compared with.  Only tests/, the
 * smoke entry point and the benchmark's CPU-baseline leg may load it; the
 * product library never links or calls it.
 *
 * Each routine follows the reference routine named in its comment, loop for
 * loop, with 0-based indices: reference irad = 1..nrad-1 is i = 0..nr-1,
 * xrad(irad) is xrad[irad] (xrad has nrad+1 entries, index 0 is r = 0).
 *
 * Third-party arithmetic the reference calls and that is restated here:
 *  - LAPACK dpbtrf/dpbtf2 + dpbtrs (band Cholesky, 'L'), netlib algorithm;
 *  - BLAS zdotu/zaxpy loops;
 *  - SHTOOLS SHGLQ (Gauss-Legendre nodes/weights, 4pi-orthonormalised
 *    associated Legendre functions with Condon-Shortley phase, complex
 *    normalisation): standard three-term recursion;
 *  - LAPACK dsyev for the Gauss-Lobatto nodes: replaced by Newton iteration
 *    on P'_n (same roots);
 *  - LAPACK zheev for the n x n Hermitian 1-RDM: cyclic complex Jacobi.
 * Parity status.  The reference cannot be built in this image and ships no
 * tests or golden vectors, so each function is pinned by closed forms,
 * library routines and invariants (tests/test_oracle_pins.py names the
 * function each test pins): GLL/Gauss-Legendre rules, Legendre tables,
 * hydrogenic spectra, dipole matrix elements, band Cholesky, the 1s Hartree
 * potential, hydrogen Slater integrals (L = 0..2, m != 0), the act2
 * generalised-Fock identity, core/active and frozen/dynamical core
 * equivalences, energy and orthonormality conservation of the equations of
 * motion, CI sigma/RDM identities, RK order, the mask closed form, the
 * dipole forms (closed forms and Ehrenfest relations), the ionisation
 * statistics (determinant generating function, moments), Gram-Schmidt and
 * imaginary-time relaxation to the He Hartree-Fock energy, and the
 * frozen-core energy (He 1s^2, all-frozen closed shell), and the CI layout
 * (determinant order by an enumeration of the ORMAS arc-weight / M-sort
 * algorithm, determinant phase by RDMs built with plain second
 * quantisation, tests/test_ci_layout.py), and the grid projection
 * (Kronecker-delta nodes, polynomial reproduction across grids, zero past
 * the source box, tests/test_wfn_proj.py).
 * Parity unpinned:
 *  - the sign convention of the m < 0 Legendre columns (copied from |m|
 *    without (-1)^m, as sph_gen.F90 does; only products of two columns enter
 *    hprod, so no pinned quantity sees it);
 *  - the KLI/LDA v2xfc (input only), and with it the type -2 part of the
 *    frozen-core energy (v2xfc enters through v2jfc and again as fcx2, as
 *    hprod_v2fc_enefc has it).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/td1c_hprod.h"

typedef double _Complex zc;

#define ORC_PI 3.14159265358979323846264338327950288

typedef struct orc {
  /* problem */
  int nfe, ndvr, nrad, nr, nradfc;
  int lmax1, lmax2, mmax1, mmax2, nlat, L1, L2;
  double znuc;
  int nfcore2, nfcore1, nfcore, ndcore, ncore, nact, nfun;
  int *mval, *lval;
  int nel[3], nelact[3], mtot;
  int icomp, igauge, isplit, ioorot, iprojfc, jfc_implicit, xfc_implicit, lconst;
  double throcc1, throcc2, rmask, mask_order;
  int nbas, nbas2, ngrid;
  /* FEDVR (clfedvr) */
  double *x0, *x1;
  double *dvr_x, *dvr_w, *dshape; /* per element: (ndvr+1), (ndvr+1)^2 */
  int *mapf, *mapb;
  double *xrad, *wrad, *mask;
  double *radk, *radk0, *radp; /* (2n+1)*(nrad+1) */
  /* angular (clsph / sph_gen) */
  double *cost, *wlat;
  double *legf1, *legb1, *legf2, *legb2;
  /* operators (clbas::gen) */
  double *alph;                 /* (lmax1+1)*(2mmax1+1) */
  zc *tmat;                     /* (2n+1)*nr*L1 */
  zc *cninv;                    /* clh1itr LU factors: (3n+1)*nr*L1, NULL until orc_h1itr_gen */
  int *cnpiv;
  zc *pmat;                     /* (2n+1)*nr */
  double *d2ll;                 /* (n+1)*nr*L2 */
  zc *zfac, *pzfac1, *pzfac2;   /* zfac,pzfac2: nr*L1*(2mmax1+1); pzfac1: L1*(2mmax1+1) */
  double *d2fac1, *d2fac2, *d2invr, *d2rpl0, *d2rpl1; /* d2fac1: L2; others nr*L2 */
  /* hprod scratch (mod_hprod) */
  zc *orb, *h0orb, *h1orb, *gorb, *v2orb;
  zc *orbg, *gorbg, *v2orbg, *torbg;
  zc *rho2, *v2sph, *v2ang;
  zc *orb0, *rhofc, *v2jfc, *v2xfc;
  zc *den1, *rden, *rrden, *int1e, *den2, *int2e;
  zc *xmat, *fmat, *bmat, *dcic;
  double ene_fcore, ene_dcore, ene_core, ene_act, ene_tot;
  /* CAS CI */
  int nstr_a, nstr_b, lcic;
  uint32_t *str_a, *str_b;
  int *det_a, *det_b;  /* [lcic] string index per determinant */
  int *det_of;         /* [nstr_b*nstr_a] -> determinant index or -1 */
  int cic0;            /* determinant ormas_cic0 sets to 1 */
  zc *ci_work;         /* nact^2 * lcic */
  /* RK work */
  zc *rk_w0, *rk_t, *rk_k;
  /* cletdrb: the clh1rat rational functions (ETD_NRAT of them) */
  int etd_ready, etd_rosen, etd_maxcyc;
  double etd_dt1, etd_thresh;
  int etd_scheme;                         /* 1 cletdrb, 2 cletdrk4 */
  double etd_dt[8];                       /* clh1rat dtime of each */
  int etd_dimn[8], etd_dimd[8];
  zc etd_rk[8];                           /* cletdrk4's CI factors (exp1rk ... f2funrk) */
  int law_ready, law_maxcyc;              /* cllawrk (h1itr_gen(dt/2) factors) */
  double law_dt, law_thresh;
  int xab_ready, xab_nstep;               /* clexpab: steps taken since gen */
  double xab_ainv[16];                    /* inverse Taylor matrix, column-major */
  zc *xab_buf;                            /* 3 K + 1 states: WfnP, hWfnP, WfnB, tWfn */
  zc etd_cn[8][4], etd_cd0[8][3], etd_cd1[8][3];
  zc etd_numer_limit[8], etd_denom_limit[8], etd_coeff_limit[8];
  zc *etd_inv[8][3];                      /* (3n+1)*nr*L1 band LU of cd0 + cd1 T_l */
  int *etd_piv[8][3];
} orc;

static void* xcalloc(size_t n, size_t sz) {
  void* p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ------------------------------------------------------------------ */
/* Gauss-Lobatto DVR (cldvr::gen, src/dvr/cldvr.cpp:57-221)            */
/* ------------------------------------------------------------------ */

/* P_n(x) with the same long double recursion as cldvr::get_legendre
 * (src/dvr/cldvr.cpp:164-190). */
static double legendre_ld(int n, double x) {
  if (n == 0) return 1.0;
  if (n == 1) return x;
  long double p0 = 1.0L, p1 = x, p2 = 0.0L;
  for (int k = 2; k <= n; ++k) {
    long double rk = (long double)k;
    long double c0 = -(rk - 1.0L) / rk;
    long double c1 = (2.0L * rk - 1.0L) / rk;
    p2 = c0 * p0 + c1 * x * p1;
    p0 = p1;
    p1 = p2;
  }
  return (double)p2;
}

/* Interior GLL nodes = roots of P'_n.  The reference takes them as
 * eigenvalues of a Jacobi matrix (cldvr::gen_xpt, :82-138, via dsyev);
 * here Newton on P'_n from Chebyshev-Lobatto starts, in long double. */
static void gll_nodes(int n, double* x) {
  x[0] = -1.0;
  x[n] = 1.0;
  for (int k = 1; k < n; ++k) {
    long double z = -cosl((long double)ORC_PI * k / n);
    for (int it = 0; it < 100; ++it) {
      /* P_n, P_{n-1} -> P'_n and P''_n */
      long double p0 = 1.0L, p1 = z;
      for (int j = 2; j <= n; ++j) {
        long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      /* p1 = P_n, p0 = P_{n-1} */
      long double dp = n * (z * p1 - p0) / (z * z - 1.0L);
      long double d2p = (2.0L * z * dp - n * (n + 1.0L) * p1) / (1.0L - z * z);
      long double dz = dp / d2p;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    x[k] = (double)z;
  }
}

static void dvr_gen(int n, double xmin, double xmax, double* xpt, double* wpt,
                    double* dshape) {
  gll_nodes(n, xpt);
  /* cldvr::gen_wpt (:140-162) */
  double xn = (double)n;
  double scale = 2.0 / (xn * (xn + 1.0));
  wpt[0] = scale;
  wpt[n] = scale;
  for (int k = 1; k < n; ++k) {
    double p = legendre_ld(n, fabs(xpt[k]));
    wpt[k] = scale / (p * p);
  }
  double lenx = 0.5 * (xmax - xmin), midx = 0.5 * (xmax + xmin);
  for (int k = 0; k <= n; ++k) {
    xpt[k] = xpt[k] * lenx + midx;
    wpt[k] = wpt[k] * lenx;
  }
  /* cldvr::gen_dshape (:192-221): dshape[(n+1)*m + k] = f_m'(x_k) */
  for (int m = 0; m <= n; ++m) {
    for (int k = 0; k <= n; ++k) {
      int km = (n + 1) * m + k;
      if (m != k) {
        double dshx = 1.0 / (xpt[m] - xpt[k]);
        for (int l = 0; l <= n; ++l) {
          if (l != m && l != k) {
            double diffm = xpt[m] - xpt[l];
            double diffk = xpt[k] - xpt[l];
            dshx *= diffk / diffm;
          }
        }
        dshape[km] = dshx;
      } else {
        dshape[km] = 0.0;
      }
    }
  }
  dshape[0] = -0.5 / wpt[0];
  dshape[(n + 1) * (n + 1) - 1] = 0.5 / wpt[n];
}

/* ------------------------------------------------------------------ */
/* FEDVR (clfedvr, src/fedvr/clfedvr.cpp)                              */
/* ------------------------------------------------------------------ */
static void fedvr_gen(orc* o) {
  int n = o->ndvr, nfe = o->nfe;
  o->nrad = n * nfe;
  int nrad = o->nrad;
  o->dvr_x = xcalloc((size_t)nfe * (n + 1), sizeof(double));
  o->dvr_w = xcalloc((size_t)nfe * (n + 1), sizeof(double));
  o->dshape = xcalloc((size_t)nfe * (n + 1) * (n + 1), sizeof(double));
  for (int i = 0; i < nfe; ++i)
    dvr_gen(n, o->x0[i], o->x1[i], o->dvr_x + i * (n + 1), o->dvr_w + i * (n + 1),
            o->dshape + (size_t)i * (n + 1) * (n + 1));
  /* gen_map (:313-338) */
  o->mapf = xcalloc(nfe, sizeof(int));
  o->mapb = xcalloc(nrad + 1, sizeof(int));
  int nb = 0;
  for (int i = 0; i < nfe; ++i) {
    o->mapf[i] = nb;
    for (int m = 0; m < n; ++m) o->mapb[nb++] = i;
  }
  o->mapb[nrad] = nfe - 1;
  /* gen_grid (:400-511), mask-type boundary (ecs off) */
  o->xrad = xcalloc(nrad + 1, sizeof(double));
  o->wrad = xcalloc(nrad + 1, sizeof(double));
  for (int ib = 0; ib <= nrad; ++ib) {
    int i = o->mapb[ib], m = ib - o->mapf[i];
    o->xrad[ib] = o->dvr_x[i * (n + 1) + m];
    o->wrad[ib] = o->dvr_w[i * (n + 1) + m];
  }
  for (int i = 1; i < nfe; ++i) o->wrad[o->mapf[i]] += o->dvr_w[(i - 1) * (n + 1) + n];
  /* gen_mask (:361-398) */
  o->mask = xcalloc(nrad + 1, sizeof(double));
  double order = o->mask_order > 0 ? 1.0 / o->mask_order : fabs(o->mask_order);
  double rul = o->x1[nfe - 1];
  for (int ib = 1; ib < nrad; ++ib) {
    int i = o->mapb[ib], m = ib - o->mapf[i];
    double xv = o->dvr_x[i * (n + 1) + m];
    double absx = sqrt(xv * xv);
    if (absx <= o->rmask) {
      o->mask[ib] = 1.0;
    } else {
      double tmp = 0.5 * ORC_PI * (absx - o->rmask) / (rul - o->rmask);
      o->mask[ib] = pow(cos(tmp), order);
    }
  }
  /* gen_radk (:676-757), fedvr_normalized */
  size_t bw = 2 * (size_t)n + 1;
  o->radk = xcalloc(bw * (nrad + 1), sizeof(double));
  o->radk0 = xcalloc(bw * (nrad + 1), sizeof(double));
  o->radp = xcalloc(bw * (nrad + 1), sizeof(double));
  for (int i = 0; i < nfe; ++i) {
    int ib0 = o->mapf[i];
    const double* w = o->dvr_w + i * (n + 1);
    const double* ds = o->dshape + (size_t)i * (n + 1) * (n + 1);
    for (int m = 0; m <= n; ++m) {
      int mb = ib0 + m;
      for (int k = 0; k <= n; ++k) {
        int kb = ib0 + k;
        size_t kmb = bw * mb + (n + kb - mb);
        if (i == 0 && (m == 0 || k == 0)) continue;
        if (i == nfe - 1 && (m == n || k == n)) continue;
        double tmp = 0.0, tmp0 = 0.0;
        for (int l = (i == 0 ? 1 : 0); l <= n; ++l)
          tmp0 += w[l] * ds[(n + 1) * m + l] * ds[(n + 1) * k + l];
        for (int l = 0; l <= n; ++l) tmp += w[l] * ds[(n + 1) * m + l] * ds[(n + 1) * k + l];
        double nrm = sqrt(o->wrad[mb] * o->wrad[kb]);
        o->radk[kmb] += 0.5 * tmp / nrm;
        o->radk0[kmb] += 0.5 * tmp0 / nrm;
      }
    }
  }
  /* gen_radp (:759-792) */
  for (int i = 0; i < nfe; ++i) {
    int ib0 = o->mapf[i];
    const double* w = o->dvr_w + i * (n + 1);
    const double* ds = o->dshape + (size_t)i * (n + 1) * (n + 1);
    for (int m = 0; m <= n; ++m) {
      int mb = ib0 + m;
      for (int k = 0; k <= n; ++k) {
        int kb = ib0 + k;
        size_t kmb = bw * mb + (n + kb - mb);
        if (i == 0 && (m == 0 || k == 0)) continue;
        if (i == nfe - 1 && (m == n || k == n)) continue;
        double tmp = w[m] * ds[(n + 1) * k + m];
        o->radp[kmb] += tmp / sqrt(o->wrad[mb] * o->wrad[kb]);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Angular tables (sph_gen, src/sph/sph_gen.F90:2-110)                 */
/* ------------------------------------------------------------------ */

/* Gauss-Legendre nodes (cos theta) and weights on [-1,1], ascending. */
static void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    long double z = cosl((long double)ORC_PI * (i + 0.75L) / (n + 0.5L));
    long double pp = 0.0L;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1.0L, p1 = z;
      for (int j = 2; j <= n; ++j) {
        long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1.0L; }
      pp = n * (z * p1 - p0) / (z * z - 1.0L);
      long double dz = p1 / pp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    /* recompute derivative at the converged root */
    {
      long double p0 = 1.0L, p1 = z;
      for (int j = 2; j <= n; ++j) {
        long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      pp = n * (z * p1 - p0) / (z * z - 1.0L);
    }
    x[i] = (double)(-z);
    x[n - 1 - i] = (double)z;
    double wi = (double)(2.0L / ((1.0L - z * z) * pp * pp));
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
}

/* sqrt(2 pi) * SHTOOLS norm=4, csphase=-1, cnorm=1 associated Legendre
 * function = sqrt((2l+1)/2 (l-m)!/(l+m)!) P_l^m(x), P_l^m with the
 * Condon-Shortley phase, for 0 <= m <= l <= lmax.  p[l] for fixed m. */
static void assoc_legendre_col(int lmax, int m, double x, double* p) {
  double s = sqrt(fmax(0.0, 1.0 - x * x));
  double pmm = sqrt(0.5);
  for (int k = 1; k <= m; ++k) pmm *= -s * sqrt((2.0 * k + 1.0) / (2.0 * k));
  for (int l = 0; l < m; ++l) p[l] = 0.0;
  if (m > lmax) return;
  p[m] = pmm;
  if (m + 1 > lmax) return;
  double pm1 = x * sqrt(2.0 * m + 3.0) * pmm;
  p[m + 1] = pm1;
  double pl2 = pmm, pl1 = pm1;
  for (int l = m + 2; l <= lmax; ++l) {
    double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
    double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
    double pl = a * (x * pl1 - b * pl2);
    p[l] = pl;
    pl2 = pl1;
    pl1 = pl;
  }
}

static void sph_gen(orc* o) {
  int nlat = o->nlat;
  o->cost = xcalloc(nlat, sizeof(double));
  o->wlat = xcalloc(nlat, sizeof(double));
  gauss_legendre(nlat, o->cost, o->wlat);
  int L1 = o->L1, L2 = o->L2, M1 = 2 * o->mmax1 + 1, M2 = 2 * o->mmax2 + 1;
  o->legf1 = xcalloc((size_t)L1 * nlat * M1, sizeof(double));
  o->legb1 = xcalloc((size_t)L1 * nlat * M1, sizeof(double));
  o->legf2 = xcalloc((size_t)L2 * nlat * M2, sizeof(double));
  o->legb2 = xcalloc((size_t)L2 * nlat * M2, sizeof(double));
  int lmx = o->lmax2 > o->lmax1 ? o->lmax2 : o->lmax1;
  double* p = xcalloc(lmx + 1, sizeof(double));
  for (int ilat = 0; ilat < nlat; ++ilat) {
    for (int m = 0; m <= o->mmax2 || m <= o->mmax1; ++m) {
      assoc_legendre_col(lmx, m, o->cost[ilat], p);
      /* negative m copies the positive-m table without a sign
       * (sph_gen.F90:85-100) */
      for (int sg = -1; sg <= 1; sg += 2) {
        int mm = sg * m;
        if (m == 0 && sg < 0) continue;
        if (m <= o->mmax1)
          for (int l = m; l <= o->lmax1; ++l) {
            o->legf1[((size_t)(mm + o->mmax1) * nlat + ilat) * L1 + l] = p[l];
            o->legb1[((size_t)(mm + o->mmax1) * L1 + l) * nlat + ilat] = p[l] * o->wlat[ilat];
          }
        if (m <= o->mmax2)
          for (int l = m; l <= o->lmax2; ++l) {
            o->legf2[((size_t)(mm + o->mmax2) * nlat + ilat) * L2 + l] = p[l];
            o->legb2[((size_t)(mm + o->mmax2) * L2 + l) * nlat + ilat] = p[l] * o->wlat[ilat];
          }
      }
    }
  }
  free(p);
}

#define LEGF1(l, ilat, m) o->legf1[((size_t)((m) + o->mmax1) * o->nlat + (ilat)) * o->L1 + (l)]
#define LEGB1(ilat, l, m) o->legb1[((size_t)((m) + o->mmax1) * o->L1 + (l)) * o->nlat + (ilat)]
#define LEGF2(l, ilat, m) o->legf2[((size_t)((m) + o->mmax2) * o->nlat + (ilat)) * o->L2 + (l)]
#define LEGB2(ilat, l, m) o->legb2[((size_t)((m) + o->mmax2) * o->L2 + (l)) * o->nlat + (ilat)]
#define ALPH(l, m) o->alph[(size_t)((m) + o->mmax1) * o->L1 + (l)]

/* ------------------------------------------------------------------ */
/* Band Cholesky (LAPACK dpbtf2/dpbtrs 'L', netlib reference algorithm) */
/* ab[(kd+1)*j + (i-j)] = A(i,j) for j <= i <= j+kd, 0-based            */
/* ------------------------------------------------------------------ */
static int band_chol(int n, int kd, double* ab) {
  int ld = kd + 1;
  for (int j = 0; j < n; ++j) {
    double ajj = ab[(size_t)ld * j];
    if (ajj <= 0.0) return j + 1;
    ajj = sqrt(ajj);
    ab[(size_t)ld * j] = ajj;
    int kn = kd < n - 1 - j ? kd : n - 1 - j;
    if (kn > 0) {
      double r = 1.0 / ajj;
      for (int i = 1; i <= kn; ++i) ab[(size_t)ld * j + i] *= r;
      /* dsyr('L', kn, -1, x, A(j+1..)) */
      for (int c = 0; c < kn; ++c) {
        double xc = ab[(size_t)ld * j + 1 + c];
        if (xc != 0.0) {
          double t = -xc;
          for (int r2 = c; r2 < kn; ++r2)
            ab[(size_t)ld * (j + 1 + c) + (r2 - c)] += ab[(size_t)ld * j + 1 + r2] * t;
        }
      }
    }
  }
  return 0;
}

static void band_chol_solve(int n, int kd, const double* ab, double* x) {
  int ld = kd + 1;
  /* dtbsv 'L','N','N' */
  for (int j = 0; j < n; ++j) {
    if (x[j] != 0.0) {
      x[j] /= ab[(size_t)ld * j];
      double t = x[j];
      int iu = j + kd < n - 1 ? j + kd : n - 1;
      for (int i = j + 1; i <= iu; ++i) x[i] -= t * ab[(size_t)ld * j + (i - j)];
    }
  }
  /* dtbsv 'L','T','N' */
  for (int j = n - 1; j >= 0; --j) {
    double t = x[j];
    int iu = j + kd < n - 1 ? j + kd : n - 1;
    for (int i = iu; i >= j + 1; --i) t -= ab[(size_t)ld * j + (i - j)] * x[i];
    t /= ab[(size_t)ld * j];
    x[j] = t;
  }
}

/* ------------------------------------------------------------------ */
/* Operator tables (clbas::gen, src/bas/clbas.cpp:100-160)             */
/* ------------------------------------------------------------------ */
static int bas_gen(orc* o) {
  int n = o->ndvr, nr = o->nr, L1 = o->L1, L2 = o->L2;
  size_t bw = 2 * (size_t)n + 1;
  int M1 = 2 * o->mmax1 + 1;
  /* bas_gen_alph (src/bas/bas_gen_alph.F90:2-27) */
  o->alph = xcalloc((size_t)L1 * M1, sizeof(double));
  for (int m = -o->mmax1; m <= o->mmax1; ++m)
    for (int l = abs(m); l <= o->lmax1; ++l) {
      double dl = l, dm = m;
      ALPH(l, m) = sqrt(((dl + 1.0) * (dl + 1.0) - dm * dm) / ((2.0 * dl + 1.0) * (2.0 * dl + 3.0)));
    }
  /* bas_gen_pmat (src/bas/bas_gen_pmat.F90), ecs off */
  o->pmat = xcalloc(bw * nr, sizeof(zc));
  for (int i = 0; i < nr; ++i) {
    int irad = i + 1;
    int jll = i - n > 0 ? i - n : 0, jul = i + n < nr - 1 ? i + n : nr - 1;
    for (int j = jll; j <= jul; ++j) {
      size_t jb = n + j - i;
      o->pmat[bw * i + jb] = o->radp[bw * irad + jb];
    }
  }
  /* bas_gen_tmat (src/bas/bas_gen_tmat.F90:2-118): pure Coulomb, ecs off */
  o->tmat = xcalloc(bw * nr * L1, sizeof(zc));
  for (int l = 0; l <= o->lmax1; ++l) {
    double l2 = (double)(l * (l + 1));
    for (int i = 0; i < nr; ++i) {
      double oor = 1.0 / o->xrad[i + 1];
      o->tmat[(bw * ((size_t)l * nr + i)) + n] = oor * (-o->znuc + l2 * oor * 0.5);
    }
  }
  for (int l = 0; l <= o->lmax1; ++l) {
    const double* rk = l == 0 ? o->radk : o->radk0;
    for (int i = 0; i < nr; ++i) {
      int jll = i - n > 0 ? i - n : 0, jul = i + n < nr - 1 ? i + n : nr - 1;
      for (int j = jll; j <= jul; ++j) {
        size_t jb = n + j - i;
        o->tmat[bw * ((size_t)l * nr + i) + jb] += rk[bw * (i + 1) + jb];
      }
    }
  }
  /* bas_gen_d2ll (src/bas/bas_gen_d2ll.F90:2-140) */
  size_t ld = n + 1;
  o->d2ll = xcalloc(ld * nr * L2, sizeof(double));
  for (int l = 0; l <= o->lmax2; ++l) {
    double l2 = (double)(l * (l + 1));
    double* ab = o->d2ll + ld * nr * l;
    for (int i = 0; i < nr; ++i) {
      double angk = 1.0 / o->xrad[i + 1];
      ab[ld * i] = angk * angk * l2;
    }
    for (int i = 0; i < nr; ++i) {
      int jul = i + n < nr - 1 ? i + n : nr - 1;
      for (int j = i; j <= jul; ++j) ab[ld * i + (j - i)] += 2.0 * o->radk[bw * (i + 1) + n + (j - i)];
    }
    int info = band_chol(nr, n, ab);
    if (info) {
      fprintf(stderr, "oracle: dpbtrf failed l=%d info=%d\n", l, info);
      return -1;
    }
  }
  /* bas_gen_zfac, bas_gen_pzfac (fedvr_normalized, ecs off) */
  o->zfac = xcalloc((size_t)nr * L1 * M1, sizeof(zc));
  o->pzfac1 = xcalloc((size_t)L1 * M1, sizeof(zc));
  o->pzfac2 = xcalloc((size_t)nr * L1 * M1, sizeof(zc));
  for (int m = -o->mmax1; m <= o->mmax1; ++m)
    for (int l = abs(m); l <= o->lmax1; ++l) {
      double a = ALPH(l, m);
      o->pzfac1[(size_t)(m + o->mmax1) * L1 + l] = -I * a;
      zc tmp = -I * a * ((double)l + 1.0);
      for (int i = 0; i < nr; ++i) {
        size_t idx = ((size_t)(m + o->mmax1) * L1 + l) * nr + i;
        o->zfac[idx] = o->xrad[i + 1] * a;
        o->pzfac2[idx] = tmp / o->xrad[i + 1];
      }
    }
  /* bas_gen_d2fac (src/bas/bas_gen_d2fac.F90:2-80) */
  o->d2fac1 = xcalloc(L2, sizeof(double));
  o->d2fac2 = xcalloc((size_t)nr * L2, sizeof(double));
  o->d2invr = xcalloc((size_t)nr * L2, sizeof(double));
  o->d2rpl0 = xcalloc((size_t)nr * L2, sizeof(double));
  o->d2rpl1 = xcalloc((size_t)nr * L2, sizeof(double));
  double r2pi = 0.5 / ORC_PI;
  for (int l = 0; l <= o->lmax2; ++l)
    for (int i = 0; i < nr; ++i) {
      double x = o->xrad[i + 1], sw = sqrt(o->wrad[i + 1]);
      size_t k = (size_t)l * nr + i;
      o->d2rpl0[k] = pow(x, l) * r2pi;
      o->d2rpl1[k] = pow(x, l + 1) * sw;
      o->d2invr[k] = (2 * l + 1) / (x * sw) * r2pi;
      o->d2fac2[k] = 4.0 * ORC_PI / (2 * l + 1) / (x * sw);
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* CAS determinant space (ormas_init with nsub = 1)                    */
/* ------------------------------------------------------------------ */
static int popcount32(uint32_t v) { return __builtin_popcount(v); }

/* Spin strings of one spin in the reference's ORMAS order, CAS (nsub = 1).
 * ormas_arcwgt_spin (src/ormas/ormas_arcwgt.F90:52-104): vertex weights
 * vtxwgt(m,k) of the (orbital, electron) graph and arc weights
 * arc(m,k) = vtxwgt(m,k) - vtxwgt(m-1,k-1) for k <= m <= k + nact - nel.
 * ormas_str_spin (src/ormas/ormas_str.F90:83-156) numbers the strings of the
 * one distribution istr = 1..C(nact,nel) and decodes each with ormas_occvec
 * (src/ormas/ormas_occvec.F90:2-29): from the highest orbital down, an
 * orbital is occupied when the remaining rank is >= its arc weight. */
static int make_strings(int nact, int nel, uint32_t** out) {
  int vtx[33][33];
  memset(vtx, 0, sizeof vtx);
  vtx[0][0] = 1;
  for (int k = 0; k <= nel; ++k) vtx[k][k] = 1;
  for (int m = 0; m <= nact - nel; ++m) vtx[m][0] = 1;
  for (int k = 1; k <= nel; ++k)
    for (int m = k + 1; m <= k + nact - nel; ++m) vtx[m][k] = vtx[m - 1][k] + vtx[m - 1][k - 1];
  int arc[33][33]; /* arc[m][k], 1-based m and k */
  memset(arc, 0, sizeof arc);
  for (int k = 1; k <= nel; ++k)
    for (int m = k; m <= k + nact - nel; ++m) arc[m][k] = vtx[m][k] - vtx[m - 1][k - 1];
  const int nstr = vtx[nact][nel];
  uint32_t* s = xcalloc(nstr > 0 ? nstr : 1, sizeof(uint32_t));
  for (int istr = 1; istr <= nstr; ++istr) {
    uint32_t v = 0;
    int iel = nel, rest = istr - 1;
    for (int iorb = nact; iorb >= 1; --iorb) {
      if (iel <= 0) break;
      if (rest >= arc[iorb][iel]) {
        v |= 1u << (iorb - 1);
        rest -= arc[iorb][iel];
        --iel;
      }
    }
    s[istr - 1] = v;
  }
  *out = s;
  return nstr;
}

/* total M of a spin string including the cores, mval_alph / mval_beta of
 * ormas_madapt_str (src/ormas/ormas_madapt.F90:170-195) */
static int str_m(const orc* o, uint32_t s) {
  int m = 0;
  for (int k = 0; k < o->ncore; ++k) m += o->mval[k];
  for (int k = 0; k < o->nact; ++k)
    if (s >> k & 1u) m += o->mval[o->ncore + k];
  return m;
}

/* a+_p a_h |s>: returns sign (0 if zero), new string in *t */
static int excite(uint32_t s, int p, int h, uint32_t* t) {
  if (!(s >> h & 1u)) return 0;
  if (p == h) { *t = s; return 1; }
  if (s >> p & 1u) return 0;
  uint32_t lo = p < h ? p : h, hi = p < h ? h : p;
  uint32_t between = s & (((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u));
  *t = (s & ~(1u << h)) | (1u << p);
  return (popcount32(between) & 1) ? -1 : 1;
}

/* ormas_madapt_str (src/ormas/ormas_madapt.F90:102-501) for one
 * distribution: the strings are reordered by their total M, keeping the
 * ormas_str order within one M (map2to1, :221-246); *first2 receives
 * map1to2(1), the new position of the first string. */
static void madapt_order(const orc* o, uint32_t* s, int n, int* first2) {
  int mmin = 0, mmax = 0;
  for (int k = 0; k < n; ++k) {
    int m = str_m(o, s[k]);
    if (m < mmin) mmin = m;
    if (m > mmax) mmax = m;
  }
  uint32_t* t = xcalloc(n, sizeof(uint32_t));
  int istr2 = 0;
  for (int m = mmin; m <= mmax; ++m)
    for (int k = 0; k < n; ++k)
      if (str_m(o, s[k]) == m) {
        if (k == 0) *first2 = istr2;
        t[istr2++] = s[k];
      }
  memcpy(s, t, n * sizeof(uint32_t));
  free(t);
}

static void ci_setup(orc* o) {
  if (o->nact == 0) {
    o->lcic = 1;
    o->cic0 = 0;
    return;
  }
  int na = o->nelact[0], nb = o->nelact[1];
  o->nstr_a = make_strings(o->nact, na, &o->str_a);
  o->nstr_b = make_strings(o->nact, nb, &o->str_b);
  int first_a = 0, first_b = 0;
  madapt_order(o, o->str_a, o->nstr_a, &first_a);
  madapt_order(o, o->str_b, o->nstr_b, &first_b);
  size_t nfull = (size_t)o->nstr_a * o->nstr_b;
  /* allowed alpha strings of each beta string (M_a + M_b = mtot, a
   * contiguous range llstr_alph_beta .. + nstr_alph_beta - 1 of the M-ordered
   * alpha strings) and the determinant offsets
   * ntot_alph_beta(b) = sum of the earlier ranges - llstr_alph_beta(b) + 1
   * (:381-435), 0-based here; determinant (a, b) is ntot_alph_beta(b) + a.
   * det_of spans the full alpha x beta product space (index ib*nstr_a + ia). */
  int* ntot = xcalloc(o->nstr_b, sizeof(int));
  o->det_of = xcalloc(nfull, sizeof(int));
  for (size_t k = 0; k < nfull; ++k) o->det_of[k] = -1;
  int cnt = 0;
  for (int ib = 0; ib < o->nstr_b; ++ib) {
    int ll = -1, nal = 0;
    for (int ia = 0; ia < o->nstr_a; ++ia)
      if (str_m(o, o->str_a[ia]) + str_m(o, o->str_b[ib]) == o->mtot) {
        if (ll < 0) ll = ia;
        ++nal;
      }
    ntot[ib] = cnt - ll; /* llstr_alph_beta = 0 (ll = -1) when none is allowed */
    for (int ia = ll; ia >= 0 && ia < ll + nal; ++ia) o->det_of[(size_t)ib * o->nstr_a + ia] = ntot[ib] + ia;
    cnt += nal;
  }
  o->lcic = cnt;
  /* ormas_cic0 (src/ormas/ormas_cic0.F90:24): the determinant of the first
   * alpha and beta strings, ntot_alph_beta(map1to2_beta(1)) + map1to2_alph(1) */
  o->cic0 = ntot[first_b] + first_a;
  free(ntot);
  o->det_a = xcalloc(cnt, sizeof(int));
  o->det_b = xcalloc(cnt, sizeof(int));
  for (int ib = 0; ib < o->nstr_b; ++ib)
    for (int ia = 0; ia < o->nstr_a; ++ia) {
      int d = o->det_of[(size_t)ib * o->nstr_a + ia];
      if (d >= 0) {
        o->det_a[d] = ia;
        o->det_b[d] = ib;
      }
    }
  o->ci_work = xcalloc((size_t)(o->nact * o->nact + 3) * nfull, sizeof(zc));
}

static int find_str(const uint32_t* s, int n, uint32_t v) {
  for (int k = 0; k < n; ++k)
    if (s[k] == v) return k;
  return -1;
}

/* y += fac * E_pq x on full-space vectors (E_pq = sum_sigma a+_p a_q) */
static void ci_apply_E(const orc* o, int p, int q, zc fac, const zc* x, zc* y) {
  int na = o->nstr_a, nb = o->nstr_b;
  for (int ib = 0; ib < nb; ++ib)
    for (int ia = 0; ia < na; ++ia) {
      zc xv = x[(size_t)ib * na + ia];
      if (xv == 0.0) continue;
      uint32_t t;
      int sg = excite(o->str_a[ia], p, q, &t);
      if (sg) y[(size_t)ib * na + find_str(o->str_a, na, t)] += fac * sg * xv;
      sg = excite(o->str_b[ib], p, q, &t);
      if (sg) y[(size_t)find_str(o->str_b, nb, t) * na + ia] += fac * sg * xv;
    }
}

static void ci_expand(const orc* o, const zc* cic, zc* full) {
  size_t nfull = (size_t)o->nstr_a * o->nstr_b;
  for (size_t f = 0; f < nfull; ++f) full[f] = o->det_of[f] >= 0 ? cic[o->det_of[f]] : 0;
}

/* ormas_mkden1 / ormas_mkden2 (src/ormas/ormas_mkden1.F90:2-30,
 * ormas_mkden2.F90:2-40): den1(a,b) = <E_ba>,
 * den2(a,b,c,d) = <a+_b a+_d a_c a_a> = <E_ba E_dc> - delta_ad <E_bc>.
 * Fortran column-major: den1[a + n*b], den2[a + n*(b + n*(c + n*d))]. */
static void ci_mkden(orc* o, const zc* cic) {
  int n = o->nact;
  if (n == 0) return;
  size_t nc = (size_t)o->nstr_a * o->nstr_b;
  zc* D = o->ci_work; /* D[(p*n+q)*nc + f] = (E_pq C) */
  zc* C = D + (size_t)n * n * nc;
  memset(D, 0, sizeof(zc) * n * n * nc);
  ci_expand(o, cic, C);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) ci_apply_E(o, p, q, 1.0, C, D + (size_t)(p * n + q) * nc);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      zc s = 0;
      const zc* e = D + (size_t)(b * n + a) * nc;
      for (size_t d = 0; d < nc; ++d) s += conj(C[d]) * e[d];
      o->den1[a + n * b] = s;
    }
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int c = 0; c < n; ++c)
        for (int dd = 0; dd < n; ++dd) {
          /* <C| E_ba E_dc |C> = (E_ab C)^+ (E_dc C) */
          const zc* u = D + (size_t)(a * n + b) * nc;
          const zc* v = D + (size_t)(dd * n + c) * nc;
          zc s = 0;
          for (size_t k = 0; k < nc; ++k) s += conj(u[k]) * v[k];
          if (a == dd) s -= o->den1[c + n * b];
          o->den2[a + n * (b + n * (c + n * dd))] = s;
        }
}

/* sig += H C with H = sum_pq int1e(p,q) E_pq
 *                   + 1/2 sum_pqrs int2e(p,q,r,s) (E_pq E_rs - delta_qr E_ps)
 * (the Hamiltonian ormas_hcic_ras applies, src/ormas/ormas_hcic_ras.F90) */
static void ci_sigma(orc* o, const zc* cic, zc* sig) {
  int n = o->nact;
  size_t nc = (size_t)o->nstr_a * o->nstr_b;
  zc* D = o->ci_work;
  zc* C = D + (size_t)n * n * nc;
  zc* S = C + nc;
  zc* tmp = S + nc;
  memset(D, 0, sizeof(zc) * n * n * nc);
  memset(S, 0, sizeof(zc) * nc);
  ci_expand(o, cic, C);
  for (int r = 0; r < n; ++r)
    for (int s = 0; s < n; ++s) ci_apply_E(o, r, s, 1.0, C, D + (size_t)(r * n + s) * nc);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      zc hpq = o->int1e[p + n * q];
      for (int r = 0; r < n; ++r) hpq -= 0.5 * o->int2e[p + n * (r + n * (r + n * q))];
      const zc* e = D + (size_t)(p * n + q) * nc;
      for (size_t d = 0; d < nc; ++d) S[d] += hpq * e[d];
    }
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      memset(tmp, 0, sizeof(zc) * nc);
      for (int r = 0; r < n; ++r)
        for (int s = 0; s < n; ++s) {
          zc g = 0.5 * o->int2e[p + n * (q + n * (r + n * s))];
          if (g == 0.0) continue;
          const zc* e = D + (size_t)(r * n + s) * nc;
          for (size_t d = 0; d < nc; ++d) tmp[d] += g * e[d];
        }
      ci_apply_E(o, p, q, 1.0, tmp, S);
    }
  for (size_t f = 0; f < nc; ++f)
    if (o->det_of[f] >= 0) sig[o->det_of[f]] += S[f];
}

/* ------------------------------------------------------------------ */
/* Hermitian eigen-decomposition (zheev stand-in, cyclic Jacobi)       */
/* a[i + n*j] Hermitian in, eigenvalues w[], eigenvectors u[:,k]        */
/* ------------------------------------------------------------------ */
static void herm_eig(int n, const zc* a_in, double* w, zc* u) {
  zc* a = xcalloc((size_t)n * n, sizeof(zc));
  memcpy(a, a_in, sizeof(zc) * n * n);
  for (int i = 0; i < n * n; ++i) u[i] = 0;
  for (int i = 0; i < n; ++i) u[i + n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += creal(a[p + n * q] * conj(a[p + n * q]));
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        zc apq = a[p + n * q];
        double mag = cabs(apq);
        if (mag < 1e-300) continue;
        double app = creal(a[p + n * p]), aqq = creal(a[q + n * q]);
        double theta = 0.5 * atan2(2.0 * mag, aqq - app);
        double c = cos(theta), s = sin(theta);
        zc ph = apq / mag; /* e^{i phi} */
        /* rotation J: columns p,q: [c, s*ph; -s*conj(ph), c] applied as A <- J^H A J */
        for (int k = 0; k < n; ++k) {
          zc akp = a[k + n * p], akq = a[k + n * q];
          a[k + n * p] = c * akp - s * conj(ph) * akq;
          a[k + n * q] = s * ph * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          zc apk = a[p + n * k], aqk = a[q + n * k];
          a[p + n * k] = c * apk - s * ph * aqk;
          a[q + n * k] = s * conj(ph) * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          zc ukp = u[k + n * p], ukq = u[k + n * q];
          u[k + n * p] = c * ukp - s * conj(ph) * ukq;
          u[k + n * q] = s * ph * ukp + c * ukq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = creal(a[i + n * i]);
  free(a);
}

/* hprod_invden_inv (src/hprod/hprod_invden.F90:62-107) */
static void invden_inv(int n, double thresh, const zc* a, zc* inva) {
  double* w = xcalloc(n, sizeof(double));
  zc* u = xcalloc((size_t)n * n, sizeof(zc));
  herm_eig(n, a, w, u);
  for (int i = 0; i < n * n; ++i) inva[i] = 0;
  for (int k = 0; k < n; ++k) {
    double diag = w[k];
    double invd = thresh > 0.0 ? diag / (diag * diag + thresh) : 1.0 / diag;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) inva[i + n * j] += u[i + n * k] * invd * conj(u[j + n * k]);
  }
  free(w);
  free(u);
}

/* ------------------------------------------------------------------ */
/* Context creation                                                    */
/* ------------------------------------------------------------------ */
orc* orc_create(const td1c_problem* p) {
  orc* o = xcalloc(1, sizeof(orc));
  o->nfe = p->nfe;
  o->ndvr = p->ndvr;
  o->x0 = xcalloc(p->nfe, sizeof(double));
  o->x1 = xcalloc(p->nfe, sizeof(double));
  for (int i = 0; i < p->nfe; ++i) {
    o->x0[i] = p->edges[i];
    o->x1[i] = p->edges[i + 1];
  }
  o->lmax1 = p->lmax1;
  o->lmax2 = p->lmax2;
  o->L1 = p->lmax1 + 1;
  o->L2 = p->lmax2 + 1;
  o->nlat = p->nlat > 0 ? p->nlat : p->lmax2 + 1;
  o->znuc = p->znuc;
  o->nfcore2 = p->nfcore2;
  o->nfcore1 = p->nfcore1;
  o->nfcore = p->nfcore2 + p->nfcore1;
  o->ndcore = p->ndcore;
  o->ncore = o->nfcore + p->ndcore;
  o->nact = p->nact;
  o->nfun = o->ncore + p->nact;
  o->mval = xcalloc(o->nfun, sizeof(int));
  o->lval = xcalloc(o->nfun, sizeof(int));
  int mmax = 0;
  for (int i = 0; i < o->nfun; ++i) {
    o->mval[i] = p->mval ? p->mval[i] : 0;
    o->lval[i] = p->lval ? p->lval[i] : 0;
    if (abs(o->mval[i]) > mmax) mmax = abs(o->mval[i]);
  }
  o->mmax1 = p->mmax1 >= 0 ? p->mmax1 : mmax;
  o->mmax2 = p->mmax2 >= 0 ? p->mmax2 : 2 * o->mmax1;
  o->nel[0] = p->nel_alpha;
  o->nel[1] = p->nel_beta;
  o->nel[2] = p->nel_alpha + p->nel_beta;
  o->nelact[0] = p->nel_alpha - o->ncore;
  o->nelact[1] = p->nel_beta - o->ncore;
  o->nelact[2] = o->nelact[0] + o->nelact[1];
  o->mtot = p->mtot;
  o->icomp = p->icomp;
  o->lconst = p->lconst;
  o->igauge = p->igauge;
  o->isplit = p->isplit;
  o->ioorot = p->ioorot;
  o->iprojfc = p->iprojfc;
  o->jfc_implicit = p->jfc_implicit;
  o->xfc_implicit = p->xfc_implicit;
  o->throcc1 = p->throcc1;
  o->throcc2 = p->throcc2;
  double R = p->edges[p->nfe];
  o->rmask = p->rmask > 0 ? p->rmask : 0.8 * R;
  o->mask_order = p->mask_cos_order > 0 ? p->mask_cos_order : 4.0;

  fedvr_gen(o);
  o->nr = o->nrad - 1;
  o->nradfc = p->nradfc > 0 ? p->nradfc : o->nr;
  sph_gen(o);
  if (bas_gen(o)) return NULL;
  ci_setup(o);

  int nr = o->nr, nf = o->nfun;
  o->nbas = nr * o->L1;
  o->nbas2 = nr * o->L2;
  o->ngrid = nr * o->nlat;
  size_t sob = (size_t)o->nbas * nf, sog = (size_t)o->ngrid * nf;
  o->orb = xcalloc(sob, sizeof(zc));
  o->h0orb = xcalloc(sob, sizeof(zc));
  o->h1orb = xcalloc(sob, sizeof(zc));
  o->gorb = xcalloc(sob, sizeof(zc));
  o->v2orb = xcalloc(sob, sizeof(zc));
  o->orb0 = xcalloc(sob, sizeof(zc));
  o->orbg = xcalloc(sog, sizeof(zc));
  o->gorbg = xcalloc(sog, sizeof(zc));
  o->v2orbg = xcalloc(sog, sizeof(zc));
  o->torbg = xcalloc(sog, sizeof(zc));
  o->rho2 = xcalloc((size_t)o->nbas2 * nf * nf, sizeof(zc));
  o->v2sph = xcalloc((size_t)o->nbas2 * nf * nf, sizeof(zc));
  o->v2ang = xcalloc((size_t)o->ngrid * nf * nf, sizeof(zc));
  o->rhofc = xcalloc(nr, sizeof(zc));
  o->v2jfc = xcalloc(nr, sizeof(zc));
  o->v2xfc = xcalloc(nr, sizeof(zc));
  int na = o->nact > 0 ? o->nact : 1;
  o->den1 = xcalloc((size_t)na * na, sizeof(zc));
  o->rden = xcalloc((size_t)na * na, sizeof(zc));
  o->rrden = xcalloc((size_t)na * na, sizeof(zc));
  o->int1e = xcalloc((size_t)na * na, sizeof(zc));
  o->den2 = xcalloc((size_t)na * na * na * na, sizeof(zc));
  o->int2e = xcalloc((size_t)na * na * na * na, sizeof(zc));
  o->xmat = xcalloc((size_t)nf * nf, sizeof(zc));
  o->fmat = xcalloc((size_t)nf * nf, sizeof(zc));
  o->bmat = xcalloc((size_t)nf * nf, sizeof(zc));
  o->dcic = xcalloc(o->lcic, sizeof(zc));
  size_t wsz = sob + o->lcic;
  o->rk_w0 = xcalloc(wsz, sizeof(zc));
  o->rk_t = xcalloc(wsz, sizeof(zc));
  o->rk_k = xcalloc(wsz * 6, sizeof(zc));
  return o;
}

void orc_destroy(orc* o) {
  if (!o) return;
  void* ptrs[] = {o->x0, o->x1, o->dvr_x, o->dvr_w, o->dshape, o->mapf, o->mapb, o->xrad, o->wrad,
                  o->mask, o->radk, o->radk0, o->radp, o->cost, o->wlat, o->legf1, o->legb1,
                  o->legf2, o->legb2, o->alph, o->tmat, o->pmat, o->d2ll, o->zfac, o->pzfac1,
                  o->pzfac2, o->d2fac1, o->d2fac2, o->d2invr, o->d2rpl0, o->d2rpl1, o->orb,
                  o->h0orb, o->h1orb, o->gorb, o->v2orb, o->orbg, o->gorbg, o->v2orbg, o->torbg,
                  o->rho2, o->v2sph, o->v2ang, o->orb0, o->rhofc, o->v2jfc, o->v2xfc, o->den1,
                  o->rden, o->rrden, o->int1e, o->den2, o->int2e, o->xmat, o->fmat, o->bmat,
                  o->dcic, o->str_a, o->str_b, o->det_a, o->det_b, o->det_of, o->ci_work,
                  o->rk_w0, o->rk_t, o->rk_k, o->mval, o->lval, o->cninv, o->cnpiv};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
  for (int r = 0; r < 8; ++r)
    for (int i = 0; i < 3; ++i) {
      free(o->etd_inv[r][i]);
      free(o->etd_piv[r][i]);
    }
  free(o->xab_buf);
  free(o);
}

/* ------------------------------------------------------------------ */
/* hprod building blocks                                               */
/* ------------------------------------------------------------------ */
#define OIDX(irad, l, ifun) (((size_t)(ifun) * o->L1 + (l)) * o->nr + (irad))
#define GIDX(irad, ilat, ifun) (((size_t)(ifun) * o->nlat + (ilat)) * o->nr + (irad))
#define SIDX2(irad, l, jfun, ifun) ((((size_t)(ifun) * o->nfun + (jfun)) * o->L2 + (l)) * o->nr + (irad))
#define GIDX2(irad, ilat, jfun, ifun) ((((size_t)(ifun) * o->nfun + (jfun)) * o->nlat + (ilat)) * o->nr + (irad))

/* bas_sph2ang1 restricted to orbitals [f0,f1) and radial rows [0,dim)
 * (src/bas/bas_sph2ang1.F90:88-156 and the _fc / _radfc variants) */
static void sph2ang1(orc* o, const zc* orb1, zc* orb2, int f0, int f1, int dim) {
  for (int ifun = f0; ifun < f1; ++ifun) {
    int mi = o->mval[ifun];
#pragma omp parallel for schedule(static)
    for (int ilat = 0; ilat < o->nlat; ++ilat) {
      zc* out = orb2 + GIDX(0, ilat, ifun);
      for (int i = 0; i < dim; ++i) out[i] = 0;
      for (int l = abs(mi); l <= o->lmax1; ++l) {
        double f = LEGF1(l, ilat, mi);
        const zc* in = orb1 + OIDX(0, l, ifun);
        for (int i = 0; i < dim; ++i) out[i] += in[i] * f;
      }
    }
  }
}

/* bas_ang2sph1 (src/bas/bas_ang2sph1.F90:88-157) */
static void ang2sph1(orc* o, const zc* orb1, zc* orb2, int f0, int f1, int dim) {
  for (int ifun = f0; ifun < f1; ++ifun) {
    int mi = o->mval[ifun];
#pragma omp parallel for schedule(static)
    for (int l = abs(mi); l <= o->lmax1; ++l) {
      zc* out = orb2 + OIDX(0, l, ifun);
      for (int i = 0; i < dim; ++i) out[i] = 0;
      for (int ilat = 0; ilat < o->nlat; ++ilat) {
        double b = LEGB1(ilat, l, mi);
        const zc* in = orb1 + GIDX(0, ilat, ifun);
        for (int i = 0; i < dim; ++i) out[i] += in[i] * b;
      }
    }
  }
}

/* hprod_tprodp (src/hprod/hprod_tprod.F90:152-235): overwrites h for
 * l >= |m|, rows [0, dim) */
static void tprod(orc* o, const zc* wfn, zc* hwfn, int f0, int f1, int dim) {
  int n = o->ndvr, nr = o->nr;
  size_t bw = 2 * (size_t)n + 1;
  for (int ifun = f0; ifun < f1; ++ifun)
#pragma omp parallel for schedule(static)
    for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
      for (int i = 0; i < dim; ++i) {
        int jll = i - n > 0 ? i - n : 0, jul = i + n < nr - 1 ? i + n : nr - 1;
        const zc* t = o->tmat + bw * ((size_t)l * nr + i);
        zc s = 0;
        for (int j = jll; j <= jul; ++j) s += wfn[OIDX(j, l, ifun)] * t[n + j - i];
        hwfn[OIDX(i, l, ifun)] = s;
      }
}

/* hprod_zprodp (src/hprod/hprod_zprod.F90:108-135) */
static void zprod(orc* o, zc zf, const zc* wfn, zc* hwfn, int f0, int f1, int dim) {
  for (int ifun = f0; ifun < f1; ++ifun) {
    int m = o->mval[ifun];
    for (int l = abs(m); l <= o->lmax1 - 1; ++l)
      for (int i = 0; i < dim; ++i) {
        zc tmp = zf * o->zfac[((size_t)(m + o->mmax1) * o->L1 + l) * o->nr + i];
        hwfn[OIDX(i, l, ifun)] += wfn[OIDX(i, l + 1, ifun)] * tmp;
        hwfn[OIDX(i, l + 1, ifun)] += wfn[OIDX(i, l, ifun)] * tmp;
      }
  }
}

/* hprod_pzprodp (src/hprod/hprod_pzprod.F90:114-213), ecs off */
static void pzprod(orc* o, zc zf, const zc* wfn, zc* hwfn, int f0, int f1, int dim) {
  int n = o->ndvr, nr = o->nr, L1 = o->L1;
  size_t bw = 2 * (size_t)n + 1;
  zc* dw = xcalloc((size_t)nr * L1, sizeof(zc));
  for (int ifun = f0; ifun < f1; ++ifun) {
    int m = o->mval[ifun];
    memset(dw, 0, sizeof(zc) * nr * L1);
    for (int l = abs(m); l <= o->lmax1; ++l)
      for (int i = 0; i < dim; ++i) {
        int jll = i - n > 0 ? i - n : 0, jul = i + n < nr - 1 ? i + n : nr - 1;
        zc s = 0;
        for (int j = jll; j <= jul; ++j) s += wfn[OIDX(j, l, ifun)] * o->pmat[bw * i + n + j - i];
        dw[(size_t)l * nr + i] += s;
      }
    for (int l = abs(m); l <= o->lmax1 - 1; ++l) {
      zc tmp1 = zf * o->pzfac1[(size_t)(m + o->mmax1) * L1 + l];
      for (int i = 0; i < dim; ++i) {
        zc tmp2 = zf * o->pzfac2[((size_t)(m + o->mmax1) * L1 + l) * nr + i];
        hwfn[OIDX(i, l, ifun)] += dw[(size_t)(l + 1) * nr + i] * tmp1 + wfn[OIDX(i, l + 1, ifun)] * tmp2;
        hwfn[OIDX(i, l + 1, ifun)] += dw[(size_t)l * nr + i] * tmp1 - wfn[OIDX(i, l, ifun)] * tmp2;
      }
    }
  }
  free(dw);
}

/* pair ranges of hprod_mkrho2p / hprod_mkv2mf_poisson / _herm:
 * type 0 = 'dyn', 1 = 'fc1dyn', 2 = 'fcfc' */
static void pair_range(orc* o, int type, int* dim, int* i0, int* i1, int* j0, int* j1) {
  if (type == 0) {
    *dim = o->nr; *i0 = o->nfcore; *i1 = o->nfun; *j0 = o->nfcore; *j1 = o->nfun;
  } else if (type == 1) {
    *dim = o->nradfc; *i0 = o->nfcore2; *i1 = o->nfcore; *j0 = o->nfcore; *j1 = o->nfun;
  } else {
    *dim = o->nr; *i0 = 0; *i1 = o->nfcore; *j0 = 0; *j1 = o->nfcore;
  }
}

/* hprod_mkrho2p (src/hprod/hprod_mkrho2.F90:115-197) followed by
 * bas_ang2sph2 of the same range (src/bas/bas_ang2sph2.F90:330-482) */
static void mkrho2(orc* o, int type) {
  int dim, i0, i1, j0, j1;
  pair_range(o, type, &dim, &i0, &i1, &j0, &j1);
  for (int ifun = i0; ifun < i1; ++ifun)
    for (int jfun = (j0 > ifun ? j0 : ifun); jfun < (j1 > ifun + 1 ? j1 : ifun + 1); ++jfun)
#pragma omp parallel for schedule(static)
      for (int ilat = 0; ilat < o->nlat; ++ilat)
        for (int i = 0; i < dim; ++i)
          o->v2ang[GIDX2(i, ilat, jfun, ifun)] =
              conj(o->orbg[GIDX(i, ilat, jfun)]) * o->orbg[GIDX(i, ilat, ifun)];
  for (int ifun = i0; ifun < i1; ++ifun)
    for (int jfun = (j0 > ifun ? j0 : ifun); jfun < (j1 > ifun + 1 ? j1 : ifun + 1); ++jfun) {
      int mji = -o->mval[jfun] + o->mval[ifun];
#pragma omp parallel for schedule(static)
      for (int l = 0; l <= o->lmax2; ++l) {
        zc* out = o->rho2 + SIDX2(0, l, jfun, ifun);
        for (int i = 0; i < dim; ++i) out[i] = 0;
        if (l < abs(mji)) continue;
        for (int ilat = 0; ilat < o->nlat; ++ilat) {
          double b = LEGB2(ilat, l, mji);
          const zc* in = o->v2ang + GIDX2(0, ilat, jfun, ifun);
          for (int i = 0; i < dim; ++i) out[i] += in[i] * b;
        }
      }
    }
}

/* hprod_mkv2mf_poissonp (src/hprod/hprod_mkv2mf.F90:118-182) */
static void poisson(orc* o, int type) {
  int dim, i0, i1, j0, j1;
  pair_range(o, type, &dim, &i0, &i1, &j0, &j1);
  int nr = o->nr, n = o->ndvr;
  for (int l = 0; l <= o->lmax2; ++l) o->d2fac1[l] = 1.0 / pow(o->xrad[dim + 1], 2 * l + 1);
  for (int ifun = i0; ifun < i1; ++ifun)
    for (int jfun = (j0 > ifun ? j0 : ifun); jfun < (j1 > ifun + 1 ? j1 : ifun + 1); ++jfun)
#pragma omp parallel
    {
      double* rr = xcalloc((size_t)dim * 2, sizeof(double));
#pragma omp for schedule(static)
      for (int l = 0; l <= o->lmax2; ++l) {
        int m = -o->mval[ifun] + o->mval[jfun];
        zc* v = o->v2sph + SIDX2(0, l, jfun, ifun);
        if (l < abs(m)) {
          for (int i = 0; i < dim; ++i) v[i] = 0;
          continue;
        }
        const zc* rho = o->rho2 + SIDX2(0, l, jfun, ifun);
        const double* rpl0 = o->d2rpl0 + (size_t)l * nr;
        const double* invr = o->d2invr + (size_t)l * nr;
        double tr = 0, ti = 0;
        for (int i = 0; i < dim; ++i) {
          rr[i] = creal(rho[i]);
          rr[dim + i] = cimag(rho[i]);
          tr += rr[i] * rpl0[i];
          ti += rr[dim + i] * rpl0[i];
          rr[i] *= invr[i];
          rr[dim + i] *= invr[i];
        }
        tr *= o->d2fac1[l];
        ti *= o->d2fac1[l];
        const double* ab = o->d2ll + (size_t)(n + 1) * nr * l;
        band_chol_solve(dim, n, ab, rr);
        band_chol_solve(dim, n, ab, rr + dim);
        const double* rpl1 = o->d2rpl1 + (size_t)l * nr;
        const double* f2 = o->d2fac2 + (size_t)l * nr;
        for (int i = 0; i < dim; ++i)
          v[i] = (rr[i] + tr * rpl1[i]) * f2[i] + (rr[dim + i] + ti * rpl1[i]) * f2[i] * I;
      }
      free(rr);
    }
}

/* bas_sph2ang2 of a pair range (src/bas/bas_sph2ang2.F90:334-463) */
static void sph2ang2(orc* o, int type) {
  int dim, i0, i1, j0, j1;
  pair_range(o, type, &dim, &i0, &i1, &j0, &j1);
  for (int ifun = i0; ifun < i1; ++ifun)
    for (int jfun = (j0 > ifun ? j0 : ifun); jfun < (j1 > ifun + 1 ? j1 : ifun + 1); ++jfun) {
      int mji = -o->mval[jfun] + o->mval[ifun];
#pragma omp parallel for schedule(static)
      for (int ilat = 0; ilat < o->nlat; ++ilat) {
        zc* out = o->v2ang + GIDX2(0, ilat, jfun, ifun);
        for (int i = 0; i < dim; ++i) out[i] = 0;
        for (int l = abs(mji); l <= o->lmax2; ++l) {
          double f = LEGF2(l, ilat, mji);
          const zc* in = o->v2sph + SIDX2(0, l, jfun, ifun);
          for (int i = 0; i < dim; ++i) out[i] += in[i] * f;
        }
      }
    }
}

/* hprod_mkv2mf_herm (src/hprod/hprod_mkv2mf.F90:257-345) */
static void v2herm(orc* o, int type) {
  int dim, i0, i1, j0, j1;
  pair_range(o, type, &dim, &i0, &i1, &j0, &j1);
  for (int ifun = i0; ifun < i1; ++ifun)
    for (int jfun = (j0 > ifun + 1 ? j0 : ifun + 1); jfun < j1; ++jfun) {
      for (int l = 0; l <= o->lmax2; ++l)
        for (int i = 0; i < dim; ++i)
          o->v2sph[SIDX2(i, l, ifun, jfun)] = conj(o->v2sph[SIDX2(i, l, jfun, ifun)]);
      for (int ilat = 0; ilat < o->nlat; ++ilat)
        for (int i = 0; i < dim; ++i)
          o->v2ang[GIDX2(i, ilat, ifun, jfun)] = conj(o->v2ang[GIDX2(i, ilat, jfun, ifun)]);
    }
}

/* hprod_mfprod2 (src/hprod/hprod_mfprod2.F90:2-34 and subroutines
 * fcx1 :158-200, dcj :238-272, dcx :273-321, actj :322-388,
 * actx :389-461, act2 :540-604), jfc_implicit = xfc_implicit = true */
static void mfprod2(orc* o) {
  int nr = o->nr, nlat = o->nlat, ng = o->ngrid;
  int nfc = o->nfcore, nfc2 = o->nfcore2, ncore = o->ncore, nf = o->nfun, na = o->nact;
  zc* v2 = o->v2ang;
#define V2G(ig, jfun, ifun) v2[((size_t)(ifun) * nf + (jfun)) * ng + (ig)]
#define WG(ig, ifun) o->orbg[(size_t)(ifun) * ng + (ig)]
  /* fcx1 */
  if (o->nfcore1 > 0) {
    int dim = o->nradfc;
    for (int ifun = nfc; ifun < nf; ++ifun)
      for (int jfun = nfc2; jfun < nfc; ++jfun)
        for (int ilat = 0; ilat < nlat; ++ilat)
          for (int i = 0; i < dim; ++i) {
            size_t ig = (size_t)ilat * nr + i;
            o->gorbg[(size_t)ifun * ng + ig] -= V2G(ig, jfun, ifun) * WG(ig, jfun);
          }
  }
  zc* d2v = xcalloc((size_t)ng * (na > 0 ? na : 1), sizeof(zc));
  /* dcj */
  if (o->ndcore > 0) {
#pragma omp parallel for schedule(static)
    for (int ig = 0; ig < ng; ++ig) d2v[ig] = 0;
    for (int jfun = nfc; jfun < ncore; ++jfun)
#pragma omp parallel for schedule(static)
      for (int ig = 0; ig < ng; ++ig) d2v[ig] += V2G(ig, jfun, jfun) * 2.0;
    for (int ifun = nfc; ifun < nf; ++ifun)
#pragma omp parallel for schedule(static)
      for (int ig = 0; ig < ng; ++ig) o->gorbg[(size_t)ifun * ng + ig] += d2v[ig] * WG(ig, ifun);
    /* dcx */
    for (int ifun = nfc; ifun < nf; ++ifun)
      for (int jfun = nfc; jfun < ncore; ++jfun)
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ng; ++ig)
          o->gorbg[(size_t)ifun * ng + ig] -= V2G(ig, jfun, ifun) * WG(ig, jfun);
  }
  if (ncore > 0 && na > 0) {
    /* actj */
#pragma omp parallel for schedule(static)
    for (int ig = 0; ig < ng; ++ig) d2v[ig] = 0;
    for (int ja = 0; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka) {
        zc dens = o->den1[ka + na * ja];
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ng; ++ig) d2v[ig] += V2G(ig, ncore + ja, ncore + ka) * dens;
      }
    for (int ifun = nfc; ifun < ncore; ++ifun)
#pragma omp parallel for schedule(static)
      for (int ig = 0; ig < ng; ++ig) o->v2orbg[(size_t)ifun * ng + ig] += d2v[ig] * WG(ig, ifun);
    /* actx */
    for (int ja = 0; ja < na; ++ja) {
      int jfun = ncore + ja;
      zc* wd = d2v + (size_t)ja * ng;
#pragma omp parallel for schedule(static)
      for (int ig = 0; ig < ng; ++ig) wd[ig] = 0;
      for (int ka = 0; ka < na; ++ka) {
        int kfun = ncore + ka;
        if (o->mval[jfun] == o->mval[kfun]) {
          zc dens = -o->den1[ka + na * ja] * 0.5;
#pragma omp parallel for schedule(static)
          for (int ig = 0; ig < ng; ++ig) wd[ig] += WG(ig, kfun) * dens;
        }
      }
      for (int ifun = nfc; ifun < ncore; ++ifun)
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ng; ++ig) o->v2orbg[(size_t)ifun * ng + ig] += wd[ig] * V2G(ig, jfun, ifun);
    }
  }
  /* act2 */
  if (na > 0 && o->nelact[2] >= 2) {
    for (size_t k = 0; k < (size_t)ng * nf; ++k) o->torbg[k] = 0;
    for (int ia = 0; ia < na; ++ia) {
      int ifun = ncore + ia;
      for (int ja = 0; ja <= ia; ++ja) {
        int jfun = ncore + ja;
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ng; ++ig) d2v[ig] = 0;
        for (int ka = 0; ka < na; ++ka) {
          int kfun = ncore + ka;
          for (int la = 0; la < na; ++la) {
            int lfun = ncore + la;
            if (o->mval[ifun] + o->mval[kfun] == o->mval[jfun] + o->mval[lfun]) {
              zc dens = o->den2[ja + na * (ia + na * (la + na * ka))];
              if (cabs(dens) > 1e-15)
#pragma omp parallel for schedule(static)
                for (int ig = 0; ig < ng; ++ig) d2v[ig] += dens * V2G(ig, kfun, lfun);
            }
          }
        }
#pragma omp parallel for schedule(static)
        for (int ig = 0; ig < ng; ++ig) o->torbg[(size_t)ifun * ng + ig] += WG(ig, jfun) * d2v[ig];
        if (ia > ja)
#pragma omp parallel for schedule(static)
          for (int ig = 0; ig < ng; ++ig) o->torbg[(size_t)jfun * ng + ig] += WG(ig, ifun) * conj(d2v[ig]);
      }
    }
    for (int ia = 0; ia < na; ++ia) {
      int ifun = ncore + ia;
      for (int ka = 0; ka < na; ++ka) {
        int kfun = ncore + ka;
        if (o->mval[ifun] == o->mval[kfun]) {
          zc tmp = o->rden[ia + na * ka];
#pragma omp parallel for schedule(static)
          for (int ig = 0; ig < ng; ++ig) o->v2orbg[(size_t)kfun * ng + ig] += tmp * o->torbg[(size_t)ifun * ng + ig];
        }
      }
    }
  }
  free(d2v);
#undef V2G
#undef WG
}

/* ------------------------------------------------------------------ */
/* hprod_htot (src/hprod/hprod_htot.F90:2-208)                          */
/* ------------------------------------------------------------------ */
void orc_htotx(orc* o, double dtime, const double* lfield, const zc* wfn, const zc* cic) {
  int nf = o->nfun, nb = o->nbas, nfc = o->nfcore, ncore = o->ncore, na = o->nact;
  zc zfield = lfield[2], efield = lfield[5];
  /* hprod_htot_init (src/hprod/hprod_htot_init.F90) */
  size_t sob = (size_t)nb * nf, sog = (size_t)o->ngrid * nf;
  memset(o->orb, 0, sob * sizeof(zc));
  memset(o->h0orb, 0, sob * sizeof(zc));
  memset(o->h1orb, 0, sob * sizeof(zc));
  memset(o->gorb, 0, sob * sizeof(zc));
  memset(o->v2orb, 0, sob * sizeof(zc));
  memset(o->orbg, 0, sog * sizeof(zc));
  memset(o->gorbg, 0, sog * sizeof(zc));
  memset(o->v2orbg, 0, sog * sizeof(zc));
  memset(o->torbg, 0, sog * sizeof(zc));
  memset(o->rho2, 0, (size_t)o->nbas2 * nf * nf * sizeof(zc));
  memset(o->v2sph, 0, (size_t)o->nbas2 * nf * nf * sizeof(zc));
  memset(o->v2ang, 0, (size_t)o->ngrid * nf * nf * sizeof(zc));
  int nn = na > 0 ? na : 1;
  memset(o->den1, 0, sizeof(zc) * nn * nn);
  memset(o->rden, 0, sizeof(zc) * nn * nn);
  memset(o->rrden, 0, sizeof(zc) * nn * nn);
  memset(o->int1e, 0, sizeof(zc) * nn * nn);
  memset(o->den2, 0, sizeof(zc) * nn * nn * nn * nn);
  memset(o->int2e, 0, sizeof(zc) * nn * nn * nn * nn);
  memset(o->xmat, 0, sizeof(zc) * nf * nf);
  memset(o->fmat, 0, sizeof(zc) * nf * nf);
  memset(o->bmat, 0, sizeof(zc) * nf * nf);
  memset(o->dcic, 0, sizeof(zc) * o->lcic);

  /* hprod_orbin + hprod_fcorb (src/hprod/hprod_orbin.F90:2-105) */
  memcpy(o->orb + (size_t)nb * nfc, wfn + (size_t)nb * nfc, sizeof(zc) * (size_t)nb * (nf - nfc));
  sph2ang1(o, o->orb, o->orbg, nfc, nf, o->nr);
  if (nfc > 0) {
    if (o->igauge == 0) {
      memcpy(o->orb, o->orb0, sizeof(zc) * (size_t)nb * nfc);
      sph2ang1(o, o->orb, o->orbg, 0, nfc, o->nradfc);
    } else {
      sph2ang1(o, o->orb0, o->orbg, 0, nfc, o->nradfc);
      for (int ilat = 0; ilat < o->nlat; ++ilat)
        for (int i = 0; i < o->nradfc; ++i) {
          zc faclv = cexp(-I * lfield[8] * o->xrad[i + 1] * o->cost[ilat]);
          for (int ifun = 0; ifun < nfc; ++ifun) o->orbg[GIDX(i, ilat, ifun)] *= faclv;
        }
      ang2sph1(o, o->orbg, o->orb, 0, nfc, o->nradfc);
    }
  }

  /* RDMs */
  if (na > 0) {
    ci_mkden(o, cic);
    /* hprod_invden */
    invden_inv(na, o->throcc1, o->den1, o->rden);
    if (ncore > 0) {
      if (na * 2 == o->nelact[2]) {
        memset(o->rrden, 0, sizeof(zc) * na * na);
      } else {
        zc* tmp = xcalloc((size_t)na * na, sizeof(zc));
        for (int k = 0; k < na * na; ++k) tmp[k] = -o->den1[k];
        for (int k = 0; k < na; ++k) tmp[k + na * k] += 2.0;
        invden_inv(na, o->throcc2, tmp, o->rrden);
        free(tmp);
      }
    }
  }

  /* atomic hamiltonian */
  tprod(o, o->orb, o->h0orb, nfc, nf, o->nr);
  if (o->iprojfc == 2) tprod(o, o->orb, o->h0orb, 0, nfc, o->nradfc);
  /* laser hamiltonian */
  if (o->igauge == 0)
    zprod(o, zfield, o->orb, o->h1orb, nfc, nf, o->nr);
  else
    pzprod(o, zfield, o->orb, o->h1orb, nfc, nf, o->nr);
  if (o->iprojfc == 2) {
    if (o->igauge == 0)
      zprod(o, zfield, o->orb, o->h1orb, 0, nfc, o->nradfc);
    else
      pzprod(o, zfield, o->orb, o->h1orb, 0, nfc, o->nradfc);
  }
  if (nfc > 0 && o->igauge == 1) zprod(o, efield, o->orb, o->h1orb, 0, nfc, o->nradfc);

  /* mean field */
  if (o->nel[2] >= 2) {
    mkrho2(o, 0);
    mkrho2(o, 1);
    poisson(o, 0);
    sph2ang2(o, 0);
    v2herm(o, 0);
    poisson(o, 1);
    sph2ang2(o, 1);
    v2herm(o, 1);
    mfprod2(o);
    ang2sph1(o, o->gorbg, o->gorb, nfc, nf, o->nr);
    ang2sph1(o, o->v2orbg, o->v2orb, nfc, nf, o->nr);
  }

  /* hprod_mkint1_sph (src/hprod/hprod_mkint1_sph.F90:2-103) */
  for (int ia = 0; ia < na; ++ia) {
    int ifun = ncore + ia;
    for (int ja = 0; ja <= ia; ++ja) {
      int jfun = ncore + ja;
      zc t = 0;
      if (o->mval[jfun] == o->mval[ifun]) {
        const zc *oj = o->orb + (size_t)jfun * nb, *g = o->gorb + (size_t)ifun * nb,
                 *h1 = o->h1orb + (size_t)ifun * nb, *h0 = o->h0orb + (size_t)ifun * nb;
        for (int k = 0; k < nb; ++k) t += conj(oj[k]) * (g[k] + h1[k] + h0[k]);
      }
      o->int1e[ja + na * ia] = t;
    }
  }
  for (int ia = 0; ia < na; ++ia)
    for (int ja = ia + 1; ja < na; ++ja) o->int1e[ja + na * ia] = conj(o->int1e[ia + na * ja]);
  /* hprod_mkint2_sph (src/hprod/hprod_mkint2_sph.F90:2-70) */
  size_t nb2 = o->nbas2;
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja <= ia; ++ja)
      for (int ka = 0; ka < na; ++ka)
        for (int la = 0; la < na; ++la) {
          int ifun = ncore + ia, jfun = ncore + ja, kfun = ncore + ka, lfun = ncore + la;
          zc t = 0;
          if (o->mval[ifun] + o->mval[kfun] == o->mval[jfun] + o->mval[lfun]) {
            const zc* r = o->rho2 + ((size_t)jfun * nf + ifun) * nb2;
            const zc* v = o->v2sph + ((size_t)lfun * nf + kfun) * nb2;
            for (size_t k = 0; k < nb2; ++k) t += r[k] * v[k];
          }
          o->int2e[ia + na * (ja + na * (ka + na * la))] = t;
        }
  for (int ia = 0; ia < na; ++ia)
    for (int ja = ia + 1; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka)
        for (int la = 0; la < na; ++la)
          o->int2e[ia + na * (ja + na * (ka + na * la))] = conj(o->int2e[ja + na * (ia + na * (la + na * ka))]);

  /* energies (hprod_ene_dcore, hprod_ene_act) */
  o->ene_dcore = 0;
  o->ene_act = 0;
  if (o->ndcore > 0) {
    zc t = 0;
    for (int ifun = nfc; ifun < ncore; ++ifun) {
      const zc *oi = o->orb + (size_t)ifun * nb, *g = o->gorb + (size_t)ifun * nb,
               *h1 = o->h1orb + (size_t)ifun * nb, *h0 = o->h0orb + (size_t)ifun * nb;
      for (int k = 0; k < nb; ++k) {
        zc h01 = h1[k] + h0[k];
        t += conj(oi[k]) * (g[k] + h01 + h01);
      }
    }
    o->ene_dcore = creal(t);
  }
  if (na > 0) {
    zc t1 = 0, t2 = 0;
    for (int ia = 0; ia < na; ++ia)
      for (int ja = 0; ja < na; ++ja) {
        t1 += o->int1e[ia + na * ja] * o->den1[ja + na * ia];
        for (int ka = 0; ka < na; ++ka)
          for (int la = 0; la < na; ++la)
            t2 += o->int2e[ia + na * (ja + na * (ka + na * la))] * o->den2[ja + na * (ia + na * (la + na * ka))];
      }
    o->ene_act = creal(t1) + creal(t2) * 0.5;
  }
  o->ene_core = o->ene_fcore + o->ene_dcore;
  o->ene_tot = o->ene_core + o->ene_act;

  zc tfac = o->icomp == 0 ? (zc)dtime : I * dtime;
#define XM(j, i) o->xmat[(size_t)(i) * nf + (j)]
#define FM(j, i) o->fmat[(size_t)(i) * nf + (j)]
  /* mkxmat_fc (src/hprod/hprod_htot_mkxmat.F90:2-77) */
  if (nfc > 0) {
    for (int ifun = nfc; ifun < nf; ++ifun)
      for (int jfun = 0; jfun < nfc; ++jfun) {
        zc t = 0;
        if (o->mval[jfun] == o->mval[ifun]) {
          const zc *h0 = o->h0orb + (size_t)jfun * nb, *h1 = o->h1orb + (size_t)jfun * nb,
                   *oi = o->orb + (size_t)ifun * nb, *oj = o->orb + (size_t)jfun * nb,
                   *g = o->gorb + (size_t)ifun * nb, *v = o->v2orb + (size_t)ifun * nb;
          for (int k = 0; k < nb; ++k) t += conj(h0[k] + h1[k]) * oi[k];
          for (int k = 0; k < nb; ++k) t += conj(oj[k]) * (g[k] + v[k]);
        }
        XM(jfun, ifun) = t * tfac;
      }
  }
  /* full h012 products <j|v2+g+h1+h0|i> */
#define H012DOT(jfun, ifun, out)                                                   \
  do {                                                                             \
    const zc *oj = o->orb + (size_t)(jfun) * nb, *v = o->v2orb + (size_t)(ifun) * nb, \
             *g = o->gorb + (size_t)(ifun) * nb, *h1 = o->h1orb + (size_t)(ifun) * nb, \
             *h0 = o->h0orb + (size_t)(ifun) * nb;                                 \
    zc acc = 0;                                                                    \
    for (int k = 0; k < nb; ++k) acc += conj(oj[k]) * (v[k] + g[k] + h1[k] + h0[k]); \
    out = acc;                                                                     \
  } while (0)
  /* mkxmat_cc (:79-166) */
  if (ncore > 0 && ncore != nfc) {
    for (int ifun = nfc; ifun < ncore; ++ifun)
      for (int jfun = nfc; jfun < ncore; ++jfun) {
        zc t = 0;
        if (o->mval[jfun] == o->mval[ifun]) H012DOT(jfun, ifun, t);
        XM(jfun, ifun) = t * tfac;
      }
  }
  /* mkxmat_ca (:549-732) / ca0 (:734-834) */
  if (na > 0 && ncore > 0 && ncore != nfc) {
    if (na * 2 == o->nelact[2]) {
      for (int ifun = nfc; ifun < ncore; ++ifun)
        for (int jfun = ncore; jfun < nf; ++jfun) {
          zc ac = 0, ca = 0;
          if (o->mval[jfun] == o->mval[ifun]) {
            H012DOT(jfun, ifun, ac);
            H012DOT(ifun, jfun, ca);
          }
          XM(jfun, ifun) = ac * tfac;
          XM(ifun, jfun) = ca * tfac;
        }
    } else {
      for (int ifun = nfc; ifun < ncore; ++ifun)
        for (int jfun = ncore; jfun < nf; ++jfun) {
          zc ac = 0, ca = 0;
          if (o->mval[jfun] == o->mval[ifun]) {
            H012DOT(jfun, ifun, ac);
            H012DOT(ifun, jfun, ca);
          }
          FM(jfun, ifun) = ac;
          FM(ifun, jfun) = ca;
        }
      for (int ifun = nfc; ifun < ncore; ++ifun)
        for (int jfun = ncore; jfun < nf; ++jfun) {
          XM(jfun, ifun) = FM(jfun, ifun) * tfac;
          XM(ifun, jfun) = FM(ifun, jfun) * tfac;
        }
      for (int ifun = nfc; ifun < ncore; ++ifun)
        for (int ja = 0; ja < na; ++ja) {
          int jfun = ncore + ja;
          FM(jfun, ifun) = FM(jfun, ifun) + FM(jfun, ifun);
          zc t = 0;
          for (int ka = 0; ka < na; ++ka) t += o->den1[ja + na * ka] * conj(FM(ifun, ncore + ka));
          FM(jfun, ifun) -= t;
        }
      for (int ifun = nfc; ifun < ncore; ++ifun)
        for (int ja = 0; ja < na; ++ja) {
          int jfun = ncore + ja;
          zc t = 0;
          for (int ka = 0; ka < na; ++ka) t += o->rrden[ja + na * ka] * FM(ncore + ka, ifun);
          o->bmat[(size_t)ifun * nf + jfun] = -tfac * t;
          o->bmat[(size_t)jfun * nf + ifun] = -conj(o->bmat[(size_t)ifun * nf + jfun]);
        }
      for (int ifun = nfc; ifun < ncore; ++ifun)
        for (int jfun = ncore; jfun < nf; ++jfun) {
          XM(jfun, ifun) += o->bmat[(size_t)ifun * nf + jfun];
          XM(ifun, jfun) += o->bmat[(size_t)jfun * nf + ifun];
        }
    }
  }
  /* mkxmat_aa1 (:228-378) with a single ORMAS subspace */
  if (na > 0) {
    for (int ia = 0; ia < na; ++ia)
      for (int ja = 0; ja < na; ++ja) {
        int ifun = ncore + ia, jfun = ncore + ja;
        if (o->mval[ifun] == o->mval[jfun]) {
          zc t;
          H012DOT(jfun, ifun, t);
          FM(jfun, ifun) = t;
          XM(jfun, ifun) = t * tfac;
        }
      }
    /* mkxmat_aa2 -> ormas_xact2 (src/ormas/ormas_xact2.F90:2-102) with
     * nsub = 1: CI sigma, phase, scaling only */
    ci_sigma(o, cic, o->dcic);
    for (int d = 0; d < o->lcic; ++d) o->dcic[d] += -o->ene_act * cic[d];
    zc fac = o->icomp == 1 ? -I * dtime : (zc)(-dtime);
    for (int d = 0; d < o->lcic; ++d) o->dcic[d] *= fac;
  }
#undef H012DOT
#undef FM
#undef XM
}

/* hprod_htot_dtorb (src/hprod/hprod_htot_dtorb.F90:2-111): the Q part is
 * (v2 + g + h1 + h0) psi for isplit = 0 (:60-67), (v2 + g + h1) psi for
 * isplit = 2 (:68-75) and (v2 + g) psi for isplit = 1 (:76-83) */
static void dtorb(orc* o, double dtime, zc* hwfn) {
  int nf = o->nfun, nb = o->nbas;
  zc tfac = o->icomp == 0 ? (zc)(-dtime) : -I * dtime;
  for (int ifun = o->nfcore; ifun < nf; ++ifun) {
    zc *h = hwfn + (size_t)ifun * nb;
    const zc *v = o->v2orb + (size_t)ifun * nb, *g = o->gorb + (size_t)ifun * nb,
             *h1 = o->h1orb + (size_t)ifun * nb, *h0 = o->h0orb + (size_t)ifun * nb;
    if (o->isplit == 0)
      for (int k = 0; k < nb; ++k) h[k] = (v[k] + g[k] + h1[k] + h0[k]) * tfac;
    else if (o->isplit == 2)
      for (int k = 0; k < nb; ++k) h[k] = (v[k] + g[k] + h1[k]) * tfac;
    else
      for (int k = 0; k < nb; ++k) h[k] = (v[k] + g[k]) * tfac;
  }
  for (int ifun = o->nfcore; ifun < nf; ++ifun)
    for (int jfun = 0; jfun < nf; ++jfun)
      if (o->mval[ifun] == o->mval[jfun]) {
        zc x = o->xmat[(size_t)ifun * nf + jfun];
        zc* h = hwfn + (size_t)ifun * nb;
        const zc* oj = o->orb + (size_t)jfun * nb;
        for (int k = 0; k < nb; ++k) h[k] += oj[k] * x;
      }
}

void orc_htot(orc* o, double dtime, const double* lfield, const zc* wfn, const zc* cic, zc* hwfn,
              zc* hcic) {
  orc_htotx(o, dtime, lfield, wfn, cic);
  memcpy(hcic, o->dcic, sizeof(zc) * o->lcic);
  memset(hwfn, 0, sizeof(zc) * (size_t)o->nbas * o->nfun);
  dtorb(o, dtime, hwfn);
  /* clhprod::htot's post-op hWfn.ladapt (src/hprod/clhprod.cpp:66-67;
   * clwfn::ladapt, src/wfn/clwfn.cpp:359-366): wfn_ladapto zeroes every
   * orbital's l != lval components (src/wfn/wfn_ladapt.F90:14-39),
   * wfn_ladapt_core those of the cores (:100-125); the CI part returns early */
  if (o->lconst == 1 || o->lconst == 2) {
    int f1 = o->lconst == 1 ? o->nfun : o->ncore;
    for (int ifun = 0; ifun < f1; ++ifun)
      for (int l = 0; l <= o->lmax1; ++l)
        if (l != o->lval[ifun])
          for (int i = 0; i < o->nr; ++i) hwfn[OIDX(i, l, ifun)] = 0;
  }
}

/* ------------------------------------------------------------------ */
/* Frozen core (hprod_set_wfn0 / hprod_v2fc, src/hprod/hprod_set_wfn0.F90) */
/* ------------------------------------------------------------------ */
void orc_set_wfn0(orc* o, const zc* wfn, const zc* cic, const zc* v2xfc_in) {
  (void)cic;
  int nb = o->nbas, nr = o->nr, nfc = o->nfcore;
  memcpy(o->orb0, wfn, sizeof(zc) * (size_t)nb * o->nfun);
  memset(o->rhofc, 0, sizeof(zc) * nr);
  memset(o->v2jfc, 0, sizeof(zc) * nr);
  memset(o->v2xfc, 0, sizeof(zc) * nr);
  o->ene_fcore = 0;
  if (nfc == 0 || o->nel[2] < 2) return;
  /* hprod_orbin with lfield = 0 (:158-165): frozen-core part of orb */
  memset(o->orb, 0, sizeof(zc) * (size_t)nb * o->nfun);
  memset(o->orbg, 0, sizeof(zc) * (size_t)o->ngrid * o->nfun);
  if (o->igauge == 0) {
    memcpy(o->orb, o->orb0, sizeof(zc) * (size_t)nb * nfc);
  } else {
    sph2ang1(o, o->orb0, o->orbg, 0, nfc, o->nradfc);
    ang2sph1(o, o->orbg, o->orb, 0, nfc, o->nradfc);
  }
  /* hprod_mkrhofc (:221-253) */
  for (int ifun = 0; ifun < nfc; ++ifun)
    if (o->mval[ifun] == 0) {
      double lfac = 2.0 * o->lval[ifun] + 1.0;
      for (int i = 0; i < o->nradfc; ++i) {
        zc c = o->orb[OIDX(i, o->lval[ifun], ifun)];
        o->rhofc[i] += conj(c) * c * lfac;
      }
    }
  /* hprod_mkv2jfc (:329-385) */
  {
    int n = o->ndvr;
    double* rr = xcalloc((size_t)nr * 2, sizeof(double));
    double fac1 = 1.0 / o->xrad[nr + 1];
    double tr = 0, ti = 0;
    for (int i = 0; i < nr; ++i) {
      rr[i] = creal(o->rhofc[i]);
      rr[nr + i] = cimag(o->rhofc[i]);
      tr += rr[i] * o->d2rpl0[i];
      ti += rr[nr + i] * o->d2rpl0[i];
      rr[i] *= o->d2invr[i];
      rr[nr + i] *= o->d2invr[i];
    }
    tr *= fac1;
    ti *= fac1;
    band_chol_solve(nr, n, o->d2ll, rr);
    band_chol_solve(nr, n, o->d2ll, rr + nr);
    for (int i = 0; i < nr; ++i)
      o->v2jfc[i] = (rr[i] + tr * o->d2rpl1[i]) * o->d2fac2[i] + (rr[nr + i] + ti * o->d2rpl1[i]) * o->d2fac2[i] * I;
    free(rr);
  }
  if (o->nfcore2 > 0 && v2xfc_in) {
    memcpy(o->v2xfc, v2xfc_in, sizeof(zc) * nr);
    if (o->xfc_implicit)
      for (int i = 0; i < nr; ++i) o->v2jfc[i] += o->v2xfc[i];
  }
  /* hprod_v2fc_enefc (:790-865), before hprod_addfc_tmat: 2 <fc|h0|fc> with
   * the bare one-body operator, plus <fc|g|fc> with g = v2jfc (fcj),
   * v2xfc on every frozen core when type -2 cores exist (fcx2) and minus
   * the exchange with the type -1 cores through the FC-FC pair potentials
   * (fcx1; hprod_mkrho2p / mkv2mf_poisson / sph2ang2 / herm 'fcfc', :167-173),
   * on the theta grid, projected back by bas_ang2sph1_fc */
  {
    int ng = o->ngrid;
    if (o->igauge == 0) sph2ang1(o, o->orb, o->orbg, 0, nfc, o->nradfc);
    if (o->nfcore1 > 0) {
      mkrho2(o, 2);
      poisson(o, 2);
      sph2ang2(o, 2);
      v2herm(o, 2);
    }
    zc* h0 = xcalloc((size_t)nb * nfc, sizeof(zc));
    tprod(o, o->orb, h0, 0, nfc, o->nradfc);
    zc t = 0;
    for (int ifun = 0; ifun < nfc; ++ifun)
      for (int l = 0; l <= o->lmax1; ++l)
        for (int i = 0; i < o->nradfc; ++i) t += conj(o->orb[OIDX(i, l, ifun)]) * h0[OIDX(i, l, ifun)];
    double enefc = 2.0 * creal(t);
    zc* g = xcalloc((size_t)ng * nfc, sizeof(zc));
    for (int ifun = 0; ifun < nfc; ++ifun)
      for (int ilat = 0; ilat < o->nlat; ++ilat)
        for (int i = 0; i < o->nradfc; ++i) {
          zc v = o->v2jfc[i] + (o->nfcore2 > 0 ? o->v2xfc[i] : 0);
          g[GIDX(i, ilat, ifun)] += v * o->orbg[GIDX(i, ilat, ifun)];
        }
    if (o->nfcore1 > 0)
      for (int ifun = 0; ifun < nfc; ++ifun)
        for (int jfun = o->nfcore2; jfun < nfc; ++jfun)
          for (int ilat = 0; ilat < o->nlat; ++ilat)
            for (int i = 0; i < o->nradfc; ++i)
              g[GIDX(i, ilat, ifun)] -= o->v2ang[GIDX2(i, ilat, jfun, ifun)] * o->orbg[GIDX(i, ilat, jfun)];
    memset(h0, 0, sizeof(zc) * (size_t)nb * nfc);
    ang2sph1(o, g, h0, 0, nfc, o->nradfc);
    t = 0;
    for (int ifun = 0; ifun < nfc; ++ifun)
      for (int l = 0; l <= o->lmax1; ++l)
        for (int i = 0; i < o->nradfc; ++i) t += conj(o->orb[OIDX(i, l, ifun)]) * h0[OIDX(i, l, ifun)];
    o->ene_fcore = enefc + creal(t);
    free(g);
    free(h0);
  }
  /* hprod_addfc_tmat (src/hprod/hprod_addfc_tmat.F90) */
  if (o->jfc_implicit) {
    size_t bw = 2 * (size_t)o->ndvr + 1;
    for (int l = 0; l <= o->lmax1; ++l)
      for (int i = 0; i < nr; ++i) o->tmat[bw * ((size_t)l * nr + i) + o->ndvr] += o->v2jfc[i];
  }
}

/* ------------------------------------------------------------------ */
/* wfn_mask (src/wfn/wfn_mask.F90:2-35)                                 */
/* ------------------------------------------------------------------ */
void orc_mask(orc* o, zc* wfn) {
  int min_irad = 1;
  for (int irad = 1; irad <= o->nr; ++irad)
    if (o->xrad[irad] > o->rmask) {
      min_irad = irad;
      break;
    }
  for (int ifun = o->nfcore; ifun < o->nfun; ++ifun)
    for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
      for (int irad = min_irad; irad <= o->nr; ++irad) wfn[OIDX(irad - 1, l, ifun)] *= o->mask[irad];
}

/* ------------------------------------------------------------------ */
/* clcrk::prop (src/prop/clcrk.cpp:55-87) with the tableaux of          */
/* set_coeff_01..04 (:183-235); lfield_stages[9*is] = field(t + c_is dt) */
/* ------------------------------------------------------------------ */
int orc_rk_coeff(int order, int* nstage, double* nodes, double* wghts, double* coeff) {
  for (int k = 0; k < 100; ++k) coeff[k] = 0;
  if (order == 1) {
    *nstage = 1; nodes[0] = 0; wghts[0] = 1;
  } else if (order == 2) {
    *nstage = 2; nodes[0] = 0; nodes[1] = 0.5; wghts[0] = 0; wghts[1] = 1; coeff[1 * 10 + 0] = 0.5;
  } else if (order == 3) {
    *nstage = 3; nodes[0] = 0; nodes[1] = 0.5; nodes[2] = 1;
    wghts[0] = 1.0 / 6.0; wghts[1] = 2.0 / 3.0; wghts[2] = 1.0 / 6.0;
    coeff[10] = 0.5; coeff[20] = -1.0; coeff[21] = 2.0;
  } else if (order == 4) {
    *nstage = 4; nodes[0] = 0; nodes[1] = 0.5; nodes[2] = 0.5; nodes[3] = 1;
    wghts[0] = 1.0 / 6.0; wghts[1] = 1.0 / 3.0; wghts[2] = 1.0 / 3.0; wghts[3] = 1.0 / 6.0;
    coeff[10] = 0.5; coeff[20] = 0; coeff[21] = 0.5; coeff[30] = 0; coeff[31] = 0; coeff[32] = 1.0;
  } else {
    return -1;
  }
  return 0;
}

int orc_rk_prop(orc* o, int order, double dtime, const double* lfield_stages, zc* wfn, int apply_mask) {
  int ns;
  double nodes[10], wghts[10], coeff[100];
  if (orc_rk_coeff(order, &ns, nodes, wghts, coeff)) return -1;
  size_t so = (size_t)o->nbas * o->nfun, sz = so + o->lcic;
  zc* w0 = o->rk_w0;
  zc* t = o->rk_t;
  memcpy(w0, wfn, sz * sizeof(zc));
  for (int is = 0; is < ns; ++is) {
    zc dtw = dtime * wghts[is];
    memcpy(t, w0, sz * sizeof(zc));
    for (int js = 0; js < is; ++js) {
      zc fac = dtime * coeff[is * 10 + js];
      const zc* k = o->rk_k + sz * js;
      for (size_t q = 0; q < sz; ++q) t[q] += fac * k[q];
    }
    zc* k = o->rk_k + sz * is;
    orc_htot(o, 1.0, lfield_stages + 9 * is, t, t + so, k, k + so);
    for (size_t q = 0; q < sz; ++q) wfn[q] += dtw * k[q];
  }
  if (apply_mask) orc_mask(o, wfn);
  return 0;
}

/* ------------------------------------------------------------------ */
/* clh1itr: the one-body half steps of split2 (src/prop/clh1itr.cpp)    */
/* ------------------------------------------------------------------ */
/* zgbtrf for kl = ku = kd (LAPACK band storage, ldab = 3 kd + 1, column j
 * at ab + ldab*j, A(i,j) at row 2 kd + i - j), restating the unblocked
 * netlib zgbtf2: partial pivoting by izamax (|re| + |im|), fill-in rows
 * zeroed, the row interchange then the rank-1 update of the trailing band. */
static int band_lu(int n, int kd, zc* ab, int* ipiv) {
  int kl = kd, ku = kd, kv = ku + kl, ld = 3 * kd + 1;
#define AB(i, j) ab[(size_t)(j) * ld + (i)]
  for (int j = ku + 1; j < (kv < n ? kv : n); ++j)
    for (int i = kv - j; i < kl; ++i) AB(i, j) = 0;
  int ju = 0;
  for (int j = 0; j < n; ++j) {
    if (j + kv < n)
      for (int i = 0; i < kl; ++i) AB(i, j + kv) = 0;
    int km = kl < n - 1 - j ? kl : n - 1 - j;
    int jp = 0;
    double best = -1.0;
    for (int i = 0; i <= km; ++i) {
      double a = fabs(creal(AB(kv + i, j))) + fabs(cimag(AB(kv + i, j)));
      if (a > best) {
        best = a;
        jp = i;
      }
    }
    ipiv[j] = jp + j;
    if (AB(kv + jp, j) == 0) return j + 1;
    int jmax = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
    if (jmax > ju) ju = jmax;
    if (jp != 0)
      for (int c = 0; c <= ju - j; ++c) {
        zc t = AB(kv + jp - c, j + c);
        AB(kv + jp - c, j + c) = AB(kv - c, j + c);
        AB(kv - c, j + c) = t;
      }
    if (km > 0) {
      zc rp = 1.0 / AB(kv, j);
      for (int i = 1; i <= km; ++i) AB(kv + i, j) *= rp;
      for (int c = 1; c <= ju - j; ++c) {
        zc y = AB(kv - c, j + c);
        if (y != 0)
          for (int i = 1; i <= km; ++i) AB(kv - c + i, j + c) -= AB(kv + i, j) * y;
      }
    }
  }
  return 0;
}

/* zgbtrs('N') for the factor of band_lu: the row interchanges with the unit
 * lower triangle (column sweeps), then the banded upper triangle (ztbsv). */
static void band_lu_solve(int n, int kd, const zc* ab, const int* ipiv, zc* b) {
  int kl = kd, ku = kd, kv = kl + ku, ld = 3 * kd + 1;
  for (int j = 0; j < n - 1; ++j) {
    int lm = kl < n - 1 - j ? kl : n - 1 - j;
    int l = ipiv[j];
    if (l != j) {
      zc t = b[l];
      b[l] = b[j];
      b[j] = t;
    }
    for (int i = 1; i <= lm; ++i) b[j + i] -= b[j] * AB(kv + i, j);
  }
  for (int j = n - 1; j >= 0; --j) {
    if (b[j] == 0) continue;
    b[j] /= AB(kv, j);
    zc t = b[j];
    int i0 = j - kv > 0 ? j - kv : 0;
    for (int i = j - 1; i >= i0; --i) b[i] -= t * AB(kv + i - j, j);
  }
#undef AB
}

/* h1itr_gen_cninv (src/prop/h1itr_gen_cninv.F90:2-58): the LU factors of
 * 1 + fac T_l, fac = i dtime/2 (icomp = 1) or dtime/2 (icomp = 0), T_l the
 * bound tmat (with v2jfc once hprod_set_wfn0 has run).  clh1itr::gen calls
 * it with dtime = dt/2 (src/prop/clh1itr.cpp:29-44). */
int orc_h1itr_gen(orc* o, double dtime) {
  int n = o->ndvr, nr = o->nr, ld = 3 * n + 1;
  size_t bw = 2 * (size_t)n + 1;
  free(o->cninv);
  free(o->cnpiv);
  o->cninv = xcalloc((size_t)ld * nr * o->L1, sizeof(zc));
  o->cnpiv = xcalloc((size_t)nr * o->L1, sizeof(int));
  zc fac = o->icomp == 1 ? I * dtime * 0.5 : (zc)(dtime * 0.5);
  for (int l = 0; l <= o->lmax1; ++l) {
    zc* ab = o->cninv + (size_t)ld * nr * l;
    for (int i = 0; i < nr; ++i) {
      ab[(size_t)ld * i + 2 * n] = 1.0;
      int jll = i - n > 0 ? i - n : 0, jul = i + n < nr - 1 ? i + n : nr - 1;
      for (int j = jll; j <= jul; ++j)
        ab[(size_t)ld * i + 2 * n + j - i] += fac * o->tmat[bw * ((size_t)l * nr + i) + n + j - i];
    }
    int info = band_lu(nr, n, ab, o->cnpiv + (size_t)nr * l);
    if (info) return -(l + 1);
  }
  return 0;
}

/* h1itr_cninv (src/prop/h1itr_cninv.F90:2-32): x <- (1 + fac T_l)^-1 x for
 * every orbital and l >= |m| */
static void h1itr_cninv(orc* o, zc* x) {
  int nr = o->nr, ld = 3 * o->ndvr + 1;
  for (int ifun = 0; ifun < o->nfun; ++ifun)
    for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
      band_lu_solve(nr, o->ndvr, o->cninv + (size_t)ld * nr * l, o->cnpiv + (size_t)nr * l,
                    x + OIDX(0, l, ifun));
}

/* h1itr_prop (src/prop/h1itr_prop.F90:2-134), iprojfc = 2 (no frozen-core
 * projection, the default of clcontrol, src/control/clcontrol.cpp:110):
 * x0 = (1 + i T dt/2)^-1 (1 - i (T [+ V]) dt/2) psi, then for isplit = 1
 * the Neumann series x_k = (1 + i T dt/2)^-1 (-i V dt/2) x_(k-1) summed while
 * the RMS orbital norm of x_(k-1) is >= thresh, at most maxcyc terms; the
 * dynamical orbitals take the sum.  Here dtime is the half step (the
 * reference's dt2 = dt/2) and V = lfield[2] z (length) or lfield[2] p_z
 * (velocity).  Returns the number of cycles run. */
int orc_h1itr_prop(orc* o, const double* lfield, double dtime, int maxcyc, double thresh, zc* wfn) {
  int nf = o->nfun, nfc = o->nfcore, nb = o->nbas;
  size_t size1 = (size_t)nb * nf;
  zc zfield = lfield[2];
  zc zfac1 = o->icomp == 1 ? -I * dtime * 0.5 : (zc)(-dtime * 0.5);
  zc zfac2 = zfac1 * zfield;
  zc* orb = xcalloc(size1, sizeof(zc));
  zc* horb = xcalloc(size1, sizeof(zc));
  zc* torb = xcalloc(size1, sizeof(zc));
  memcpy(orb, wfn, size1 * sizeof(zc));
  memcpy(torb, wfn, size1 * sizeof(zc));
  tprod(o, orb, horb, nfc, nf, o->nr);
  if (o->isplit == 1) {
    if (o->igauge == 0)
      zprod(o, zfield, orb, horb, nfc, nf, o->nr);
    else
      pzprod(o, zfield, orb, horb, nfc, nf, o->nr);
  }
  for (size_t k = (size_t)nb * nfc; k < size1; ++k) orb[k] += zfac1 * horb[k];
  h1itr_cninv(o, orb);
  memcpy(torb, orb, size1 * sizeof(zc));
  int ncyc = 0;
  if (o->isplit == 1) {
    for (int icyc = 1; icyc <= maxcyc; ++icyc) {
      ncyc = icyc;
      double rms = 0.0;
      for (int ifun = 0; ifun < nf; ++ifun) {
        double t = 0.0;
        const zc* x = orb + (size_t)ifun * nb;
        for (int k = 0; k < nb; ++k) t += creal(x[k]) * creal(x[k]) + cimag(x[k]) * cimag(x[k]);
        rms += t;
      }
      rms = sqrt(rms / nf);
      if (rms < thresh) break;
      memset(horb, 0, size1 * sizeof(zc));
      if (o->igauge == 0)
        zprod(o, zfac2, orb, horb, nfc, nf, o->nr);
      else
        pzprod(o, zfac2, orb, horb, nfc, nf, o->nr);
      memcpy(orb, horb, size1 * sizeof(zc));
      h1itr_cninv(o, orb);
      for (size_t k = 0; k < size1; ++k) torb[k] += orb[k];
    }
  }
  memcpy(wfn + (size_t)nb * nfc, torb + (size_t)nb * nfc, sizeof(zc) * (size1 - (size_t)nb * nfc));
  free(orb);
  free(horb);
  free(torb);
  return ncyc;
}

/* One step of tdse_split2 with clh1itr + clcrk (src/main/tdse.cpp:225-266):
 * H1P.prop(STEP1) with the field at t + dt/4, the RK step over dt with the
 * split hprod (stage fields lfield_stages), H1P.prop(STEP2) with the field
 * at t + 3 dt/4 (clh1itr::prop, src/prop/clh1itr.cpp:61-91), then the mask.
 * h1itr_gen must have been called with dt/2. */
int orc_split2_step(orc* o, int order, double dtime, const double* lf_quarter, const double* lfield_stages,
                    const double* lf_three_quarter, int maxcyc, double thresh, zc* wfn, int apply_mask,
                    int* ncyc) {
  ncyc[0] = orc_h1itr_prop(o, lf_quarter, 0.5 * dtime, maxcyc, thresh, wfn);
  int rc = orc_rk_prop(o, order, dtime, lfield_stages, wfn, 0);
  if (rc) return rc;
  ncyc[1] = orc_h1itr_prop(o, lf_three_quarter, 0.5 * dtime, maxcyc, thresh, wfn);
  if (apply_mask) orc_mask(o, wfn);
  return 0;
}

/* ------------------------------------------------------------------ */
/* cletdrb: exponential Rosenbrock RK4 with Pade phi functions          */
/* (src/prop/cletdrb.cpp, clh1rat src/prop/clh1rat.cpp, h1rat_gen /      */
/* h1rat_prod src/prop/h1rat_gen.F90, h1rat_prod.F90), with the defaults */
/* of cletdrb::gen (:25-92): etd_rk_order 4, numerator / denominator     */
/* degrees 2 / 3 for phi1, phi2, phi3, etd_rosenbrock as given,          */
/* etd_time_of_h 0, etd_cisplit -1, split_type 1.                       */
/* ------------------------------------------------------------------ */
enum { ETD_PHI1_DT1, ETD_PHI1_DT2, ETD_PHID1_DT1, ETD_PHID1_DT2, ETD_PHI2_DT1, ETD_PHI2_DT2, ETD_PHI3_DT1, ETD_NRAT };

/* Numerator coefficients, sign of the first denominator factor and the
 * roots of the denominator of the [2/3] forms in gen_phi1 (:636-771),
 * gen_phid1 (:773-886), gen_phi2 (:888-1035), gen_phi3 (:1037-1185). */
static void etd_family(int fam, double* dfac, double* n, zc* roots) {
  static const double r1[3][2] = {{4.644370709252173768, 0.0},
                                  {3.677814645373918445, 3.508761919567442433},
                                  {3.677814645373918445, -3.508761919567442433}};
  static const double r2[3][2] = {{5.648485971016884655, 0.0},
                                  {4.675757014491553676, 3.913489560603714335},
                                  {4.675757014491553676, -3.913489560603714335}};
  static const double r3[3][2] = {{6.651316808950019421, 0.0},
                                  {5.674341595524996507, 4.279971984629763249},
                                  {5.674341595524996507, -4.279971984629763249}};
  const double(*r)[2] = fam <= 1 ? r1 : (fam == 2 ? r2 : r3);
  *dfac = -1.0;
  if (fam == 0) { n[0] = 120.0; n[1] = 0.0; n[2] = 2.0; }
  else if (fam == 1) { n[0] = 60.0; n[1] = -10.0; n[2] = 1.0; }
  else if (fam == 2) { n[0] = 420.0 / 4.0; n[1] = -40.0 / 4.0; n[2] = 5.0 / 4.0; }
  else { n[0] = 1120.0 / 20.0; n[1] = -140.0 / 20.0; n[2] = 11.0 / 20.0; }
  for (int i = 0; i < 3; ++i) roots[i] = r[i][0] + I * r[i][1];
}

/* One clh1rat from a form (numerator n[0..dimn], first-factor sign dfac,
 * denominator roots[0..dimd-1]): the scaling of the gen_* loops (numerator
 * by nscale tunit^k, denominator slopes by tunit dt) and clh1rat::gen
 * (:15-47) with h1rat_gen's band LU of cd0 + cd1 T_l (T = tmat), skipped for
 * |cd1| <= h1rat_thresh. */
static int etd_rat_set(orc* o, int ir, double dfac, const double* nn, int dimn, const zc* roots, int dimd,
                       double nscale, double dt) {
  zc tunit = o->icomp == 0 ? -1.0 : -I;
  zc tfac = tunit * dt, ttmp = nscale;
  for (int k = 0; k <= dimn; ++k) {
    o->etd_cn[ir][k] = nn[k] * ttmp;
    ttmp *= tunit;
  }
  o->etd_cd1[ir][0] = dfac * tfac;
  o->etd_cd0[ir][0] = -dfac * roots[0];
  for (int i = 1; i < dimd; ++i) {
    o->etd_cd1[ir][i] = tfac;
    o->etd_cd0[ir][i] = -roots[i];
  }
  o->etd_dimn[ir] = dimn;
  o->etd_dimd[ir] = dimd;
  o->etd_dt[ir] = dt;
  o->etd_numer_limit[ir] = o->etd_cn[ir][0];
  zc den = 1.0;
  for (int i = 0; i < dimd; ++i) den *= o->etd_cd0[ir][i];
  o->etd_denom_limit[ir] = den;
  o->etd_coeff_limit[ir] = o->etd_numer_limit[ir] / den;
  int n = o->ndvr, nr = o->nr, ld = 3 * n + 1;
  size_t bw = 2 * (size_t)n + 1;
  for (int i = 0; i < 3; ++i) {
    free(o->etd_inv[ir][i]);
    free(o->etd_piv[ir][i]);
    o->etd_inv[ir][i] = NULL;
    o->etd_piv[ir][i] = NULL;
    if (i >= dimd || !(cabs(o->etd_cd1[ir][i]) > o->etd_thresh)) continue;
    o->etd_inv[ir][i] = xcalloc((size_t)ld * nr * o->L1, sizeof(zc));
    o->etd_piv[ir][i] = xcalloc((size_t)nr * o->L1, sizeof(int));
    for (int l = 0; l <= o->lmax1; ++l) {
      zc* ab = o->etd_inv[ir][i] + (size_t)ld * nr * l;
      for (int r = 0; r < nr; ++r) {
        ab[(size_t)ld * r + 2 * n] = o->etd_cd0[ir][i];
        int jll = r - n > 0 ? r - n : 0, jul = r + n < nr - 1 ? r + n : nr - 1;
        for (int j = jll; j <= jul; ++j)
          ab[(size_t)ld * r + 2 * n + j - r] += o->tmat[bw * ((size_t)l * nr + r) + n + j - r] * o->etd_cd1[ir][i];
      }
      if (band_lu(nr, n, ab, o->etd_piv[ir][i] + (size_t)nr * l)) return -1;
    }
  }
  return 0;
}

/* cletdrb's forms: the scaling of gen_phi* (numerator by dt1 for both step
 * sizes). */
static int etd_rat_gen(orc* o, int ir, int fam, double dt) {
  double dfac, nn[3];
  zc roots[3];
  etd_family(fam, &dfac, nn, roots);
  return etd_rat_set(o, ir, dfac, nn, 2, roots, 3, o->etd_dt1, dt);
}

/* cletdrb::gen: the rational functions prop4_orb uses, for step dtime; the
 * band factors take tmat as it is now (with v2jfc after orc_set_wfn0). */
int orc_etdrb_gen(orc* o, double dtime, int rosenbrock, int maxcyc, double thresh) {
  if (o->isplit != 1) return -2;
  o->etd_dt1 = dtime;
  o->etd_rosen = rosenbrock;
  o->etd_maxcyc = maxcyc;
  o->etd_thresh = thresh;
  const int fam[ETD_NRAT] = {0, 0, 1, 1, 2, 2, 3};
  const double dts[ETD_NRAT] = {dtime, 0.5 * dtime, dtime, 0.5 * dtime, dtime, 0.5 * dtime, dtime};
  for (int r = 0; r < ETD_NRAT; ++r)
    if (etd_rat_gen(o, r, fam[r], dts[r])) return -1;
  o->etd_scheme = 1;
  o->etd_ready = 1;
  return 0;
}

/* The scalar value of rational function ir at an eigenvalue lam of h1:
 * sum_k cn[k] (dt lam)^k / prod_i (cd0[i] + cd1[i] lam), the form
 * h1rat_prod applies to orbitals (for tests of the tables). */
void orc_etd_rat_scalar(const orc* o, int ir, double lam_re, double lam_im, double* out) {
  zc lam = lam_re + I * lam_im, x = o->etd_dt[ir] * lam, num = 0.0, p = 1.0;
  for (int k = 0; k <= o->etd_dimn[ir]; ++k) {
    num += o->etd_cn[ir][k] * p;
    p *= x;
  }
  zc den = 1.0;
  for (int i = 0; i < o->etd_dimd[ir]; ++i) den *= o->etd_cd0[ir][i] + o->etd_cd1[ir][i] * lam;
  zc v = num / den;
  out[0] = creal(v);
  out[1] = cimag(v);
}

/* hprod_h1tot with dofc = false (src/hprod/hprod_h1tot.F90:2-53):
 * hw = (T + V(lf)) w on the dynamical orbitals, the rest of hw cleared (as
 * clhprod::h1tot clears it, src/hprod/clhprod.cpp:110-118). */
static void etd_h1tot(orc* o, const double* lf, const zc* w, zc* hw) {
  memset(hw, 0, sizeof(zc) * ((size_t)o->nbas * o->nfun + o->lcic));
  tprod(o, w, hw, o->nfcore, o->nfun, o->nr);
  if (o->igauge == 0)
    zprod(o, lf[2], w, hw, o->nfcore, o->nfun, o->nr);
  else
    pzprod(o, lf[2], w, hw, o->nfcore, o->nfun, o->nr);
}

/* hprod_v1ext with dofc = false (src/hprod/hprod_v1ext.F90:2-36):
 * hw += V(zfac lf) w on the dynamical orbitals. */
static void etd_v1ext(orc* o, zc zfac, const double* lf, const zc* w, zc* hw) {
  zc zf = lf[2] * zfac;
  if (o->igauge == 0)
    zprod(o, zf, w, hw, o->nfcore, o->nfun, o->nr);
  else
    pzprod(o, zf, w, hw, o->nfcore, o->nfun, o->nr);
}

/* h1rat_h0inv_v2 (h1rat_prod.F90): x <- (cd0 + cd1 T_l)^-1 x for every
 * dynamical orbital and l. */
static void etd_h0inv(orc* o, int ir, int i, zc* x) {
  int nr = o->nr, ld = 3 * o->ndvr + 1;
  for (int ifun = o->nfcore; ifun < o->nfun; ++ifun)
    for (int l = 0; l <= o->lmax1; ++l)
      band_lu_solve(nr, o->ndvr, o->etd_inv[ir][i] + (size_t)ld * nr * l, o->etd_piv[ir][i] + (size_t)nr * l,
                    x + OIDX(0, l, ifun));
}

/* h1rat_prod (h1rat_prod.F90:2-121) on the orbitals of w, field lf. */
static void etd_h1rat_prod(orc* o, int ir, int do_numer, int do_denom, const double* lf, zc* w) {
  size_t nb = o->nbas, size1 = nb * o->nfun, d0 = nb * o->nfcore;
  zc* orb = xcalloc(size1, sizeof(zc));
  zc* torb = xcalloc(size1, sizeof(zc));
  zc* h1orb = xcalloc(size1 + o->lcic, sizeof(zc));
  double tfac = o->etd_dt[ir];
  memcpy(orb + d0, w + d0, sizeof(zc) * (size1 - d0));
  if (do_numer) {
    memcpy(torb + d0, w + d0, sizeof(zc) * (size1 - d0));
    for (size_t k = d0; k < size1; ++k) orb[k] *= o->etd_cn[ir][0];
    for (int idim = 1; idim <= o->etd_dimn[ir]; ++idim) {
      etd_h1tot(o, lf, torb, h1orb);
      for (size_t k = d0; k < size1; ++k) torb[k] = tfac * h1orb[k];
      for (size_t k = d0; k < size1; ++k) orb[k] += o->etd_cn[ir][idim] * torb[k];
    }
  }
  if (do_denom) {
    for (int idim = 0; idim < o->etd_dimd[ir]; ++idim) {
      if (cabs(o->etd_cd1[ir][idim]) < o->etd_thresh) continue;
      etd_h0inv(o, ir, idim, orb);
      if (o->isplit == 1 && o->icomp == 1) {
        memcpy(torb + d0, orb + d0, sizeof(zc) * (size1 - d0));
        for (int icyc = 1; icyc <= o->etd_maxcyc; ++icyc) {
          double rms = 0.0;
          for (int ifun = o->nfcore; ifun < o->nfun; ++ifun) {
            double t = 0.0;
            const zc* x = orb + (size_t)ifun * nb;
            for (size_t k = 0; k < nb; ++k) t += creal(x[k]) * creal(x[k]) + cimag(x[k]) * cimag(x[k]);
            rms += t;
          }
          rms = sqrt(rms / o->nfun);
          if (rms < o->etd_thresh) break;
          memset(h1orb, 0, sizeof(zc) * size1);
          etd_v1ext(o, -o->etd_cd1[ir][idim], lf, orb, h1orb);
          memcpy(orb + d0, h1orb + d0, sizeof(zc) * (size1 - d0));
          etd_h0inv(o, ir, idim, orb);
          for (size_t k = d0; k < size1; ++k) torb[k] += orb[k];
        }
        memcpy(orb + d0, torb + d0, sizeof(zc) * (size1 - d0));
      }
    }
  }
  memcpy(w + d0, orb + d0, sizeof(zc) * (size1 - d0));
  free(orb);
  free(torb);
  free(h1orb);
}

/* clh1rat::prod_numer (clh1rat.cpp:196-216) with h1rat_prod_numer
 * (h1rat_prod.F90:123-181): wout.ci += zfac numer_limit win.ci,
 * wout.orb(dyn) += zfac N(dt h1) win. */
static void etd_prod_numer(orc* o, int ir, zc zfac, const double* lf, const zc* win, zc* wout) {
  size_t nb = o->nbas, size1 = nb * o->nfun, d0 = nb * o->nfcore;
  for (int k = 0; k < o->lcic; ++k) wout[size1 + k] += zfac * o->etd_numer_limit[ir] * win[size1 + k];
  zc* orb = xcalloc(size1, sizeof(zc));
  zc* torb = xcalloc(size1, sizeof(zc));
  zc* h1orb = xcalloc(size1 + o->lcic, sizeof(zc));
  double tfac = o->etd_dt[ir];
  for (size_t k = d0; k < size1; ++k) {
    orb[k] = win[k];
    torb[k] = win[k];
    orb[k] *= o->etd_cn[ir][0];
  }
  for (int idim = 1; idim <= o->etd_dimn[ir]; ++idim) {
    etd_h1tot(o, lf, torb, h1orb);
    for (size_t k = d0; k < size1; ++k) torb[k] = tfac * h1orb[k];
    for (size_t k = d0; k < size1; ++k) orb[k] += o->etd_cn[ir][idim] * torb[k];
  }
  for (size_t k = d0; k < size1; ++k) wout[k] += zfac * orb[k];
  free(orb);
  free(torb);
  free(h1orb);
}

/* clh1rat::prod_denom (:360-384) / prod2 (:108-131): wout.ci += c win.ci
 * (c = zfac / denom_limit, or zfac coeff_limit), win.orb <- D(h) win or
 * R(h) win in place, wout.orb += zfac win.orb. */
static void etd_prod(orc* o, int ir, int numer, zc zfac, const double* lf, zc* win, zc* wout) {
  size_t size1 = (size_t)o->nbas * o->nfun;
  zc c = numer ? zfac * o->etd_coeff_limit[ir] : zfac / o->etd_denom_limit[ir];
  for (int k = 0; k < o->lcic; ++k) wout[size1 + k] += c * win[size1 + k];
  etd_h1rat_prod(o, ir, numer, 1, lf, win);
  for (size_t k = 0; k < size1; ++k) wout[k] += zfac * win[k];
}

/* cletdrb::prop4_orb (src/prop/cletdrb.cpp:321-434), then the mask of
 * tdse_prop12 (src/main/tdse.cpp:268-309).  lf holds six field arrays of 9:
 * lfield(t), lfield(t + dt/2), lfield(t + dt), d lfield/dt at t
 * (clfield::get_der), lfield(t + dt/2) - lfield(t), lfield(t + dt) -
 * lfield(t).  wfn: orbitals then CI. */
int orc_etdrb_step(orc* o, const double* lf, zc* wfn, int apply_mask) {
  if (!o->etd_ready || o->etd_scheme != 1) return -1;
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  const double *lf0 = lf, *lf1 = lf + 9, *lf3 = lf + 18, *dlf0 = lf + 27, *dv1 = lf + 36, *dv3 = lf + 45;
  zc tunit = o->icomp == 0 ? -1.0 : -I;
  double dt1 = o->etd_dt1;
  zc* buf = xcalloc(sz * 10, sizeof(zc));
  zc *w0 = buf, *w1 = buf + sz, *w2 = buf + 2 * sz, *w3 = buf + 3 * sz, *hw0 = buf + 4 * sz, *hw1 = buf + 5 * sz,
     *hw2 = buf + 6 * sz, *hw3 = buf + 7 * sz, *tw = buf + 8 * sz, *dw = buf + 9 * sz;
  zc c1 = 0.5, c3 = 1.0;
  memcpy(w0, wfn, sizeof(zc) * sz);
  /* stage 0 */
  etd_h1tot(o, lf0, w0, tw);
  orc_htot(o, 1.0, lf0, w0, w0 + size1, hw0, hw0 + size1);
  for (size_t k = 0; k < sz; ++k) tw[k] = tunit * tw[k] + hw0[k];
  etd_prod_numer(o, ETD_PHI1_DT2, c1, lf0, tw, w1);
  etd_prod_numer(o, ETD_PHI1_DT1, c3, lf0, tw, w3);
  if (o->etd_rosen) {
    memset(dw, 0, sizeof(zc) * sz);
    etd_v1ext(o, 1.0, dlf0, w0, dw);
    for (size_t k = 0; k < sz; ++k) dw[k] *= tunit * dt1;
    etd_prod_numer(o, ETD_PHID1_DT2, c1 * c1, lf0, dw, w1);
    etd_prod_numer(o, ETD_PHID1_DT1, c3 * c3, lf0, dw, w3);
  }
  /* stage 1 */
  memcpy(tw, w1, sizeof(zc) * sz);
  memcpy(w1, w0, sizeof(zc) * sz);
  etd_prod(o, ETD_PHI1_DT2, 0, 1.0, lf0, tw, w1);
  orc_htot(o, 1.0, lf1, w1, w1 + size1, hw1, hw1 + size1);
  for (size_t k = 0; k < sz; ++k) hw1[k] -= hw0[k];
  if (o->etd_rosen) {
    etd_v1ext(o, tunit, dv1, w1, hw1);
    for (size_t k = 0; k < sz; ++k) hw1[k] += -0.5 * dw[k];
  }
  /* stage 2 */
  memcpy(w2, w1, sizeof(zc) * sz);
  memcpy(tw, hw1, sizeof(zc) * sz);
  etd_prod(o, ETD_PHI2_DT2, 1, 1.0, lf0, tw, w2);
  orc_htot(o, 1.0, lf1, w2, w2 + size1, hw2, hw2 + size1);
  for (size_t k = 0; k < sz; ++k) hw2[k] -= hw0[k];
  if (o->etd_rosen) {
    etd_v1ext(o, tunit, dv1, w2, hw2);
    for (size_t k = 0; k < sz; ++k) hw2[k] += -0.5 * dw[k];
  }
  /* stage 3 */
  memcpy(tw, w3, sizeof(zc) * sz);
  memcpy(w3, w0, sizeof(zc) * sz);
  etd_prod(o, ETD_PHI1_DT1, 0, 1.0, lf0, tw, w3);
  memcpy(tw, hw2, sizeof(zc) * sz);
  etd_prod(o, ETD_PHI2_DT1, 1, 2.0, lf0, tw, w3);
  orc_htot(o, 1.0, lf3, w3, w3 + size1, hw3, hw3 + size1);
  for (size_t k = 0; k < sz; ++k) hw3[k] -= hw0[k];
  if (o->etd_rosen) {
    etd_v1ext(o, tunit, dv3, w3, hw3);
    for (size_t k = 0; k < sz; ++k) hw3[k] += -1.0 * dw[k];
  }
  /* summation */
  memcpy(wfn, w3, sizeof(zc) * sz);
  for (size_t k = 0; k < sz; ++k) tw[k] = 2.0 * hw1[k] - hw3[k];
  etd_prod(o, ETD_PHI2_DT1, 1, 1.0, lf0, tw, wfn);
  for (size_t k = 0; k < sz; ++k) tw[k] = hw1[k] + hw2[k];
  for (size_t k = 0; k < sz; ++k) tw[k] = tw[k] - hw3[k];
  etd_prod(o, ETD_PHI3_DT1, 1, -4.0, lf0, tw, wfn);
  if (apply_mask) orc_mask(o, wfn);
  free(buf);
  return 0;
}

/* ------------------------------------------------------------------ */
/* cletdrk4 (src/prop/cletdrk4.cpp): the Cox-Matthews exponential RK4   */
/* with the [2/2] forms of etd_dim_numer = etd_dim_denom = 2 (gen :24-60, */
/* the defaults) for exp, gfun (phi1), f0fun, f1fun, f2fun.              */
/* ------------------------------------------------------------------ */
enum { ETK_EXP1, ETK_EXP2, ETK_GFUN, ETK_F0, ETK_F1, ETK_F2, ETK_NRAT };

/* cletdrk4::gen with gen_exp (:134-276), gen_gfun (:278-439), gen_f0fun
 * (:441-604), gen_f1fun (:606-768), gen_f2fun (:770-933). */
int orc_etdrk4_gen(orc* o, double dtime, int maxcyc, double thresh) {
  if (o->isplit != 1) return -2;
  double dt1 = dtime, dt2 = 0.5 * dtime, s3 = sqrt(3.0);
  o->etd_dt1 = dtime;
  o->etd_rosen = 0;
  o->etd_maxcyc = maxcyc;
  o->etd_thresh = thresh;
  const double n_exp[3] = {12.0, 6.0, 1.0}, n_g[3] = {60.0, 6.0, 1.0}, n_f0[3] = {34860.0, 19620.0, 2657.0},
               n_f1[3] = {2940.0, 450.0, 41.0}, n_f2[3] = {1260.0, -420.0, -13.0};
  const zc r_exp[2] = {3.0 - s3 * I, 3.0 + s3 * I}, r_g[2] = {4.0 - 2.0 * I, 4.0 + 2.0 * I},
           r_f0[2] = {(762.0 - sqrt(189762.0) * I) / 221.0, (762.0 + sqrt(189762.0) * I) / 221.0},
           r_f1[2] = {(51.0 - sqrt(633.0) * I) / 11.0, (51.0 + sqrt(633.0) * I) / 11.0},
           r_f2[2] = {(21.0 - 3.0 * sqrt(21.0) * I) / 5.0, (21.0 + 3.0 * sqrt(21.0) * I) / 5.0};
  if (etd_rat_set(o, ETK_EXP1, 1.0, n_exp, 2, r_exp, 2, 1.0, dt1) ||
      etd_rat_set(o, ETK_EXP2, 1.0, n_exp, 2, r_exp, 2, 1.0, dt2) ||
      etd_rat_set(o, ETK_GFUN, 3.0, n_g, 2, r_g, 2, dt2, dt2) ||
      etd_rat_set(o, ETK_F0, 13260.0, n_f0, 2, r_f0, 2, dt1, dt1) ||
      etd_rat_set(o, ETK_F1, 330.0, n_f1, 2, r_f1, 2, dt1, dt1) ||
      etd_rat_set(o, ETK_F2, 300.0, n_f2, 2, r_f2, 2, dt1, dt1))
    return -1;
  o->etd_rk[ETK_EXP1] = 1.0;
  o->etd_rk[ETK_EXP2] = 1.0;
  o->etd_rk[ETK_GFUN] = dt2;
  o->etd_rk[ETK_F0] = dt1 / 6.0;
  o->etd_rk[ETK_F1] = dt1 / 3.0;
  o->etd_rk[ETK_F2] = dt1 / 6.0;
  o->etd_scheme = 2;
  o->etd_ready = 1;
  return 0;
}

/* cletdrk4::prod1 (:116-123): wout = win with its CI times rk and its
 * orbitals through R (clh1rat::prod, :49-65). */
static void etk_prod1(orc* o, int ir, const double* lf, const zc* win, zc* wout) {
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  memcpy(wout, win, sizeof(zc) * sz);
  for (int k = 0; k < o->lcic; ++k) wout[size1 + k] *= o->etd_rk[ir];
  etd_h1rat_prod(o, ir, 1, 1, lf, wout);
}

/* cletdrk4::prod2 (:125-132): win's CI times rk, its orbitals through R,
 * then wout += win. */
static void etk_prod2(orc* o, int ir, const double* lf, zc* win, zc* wout) {
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  for (int k = 0; k < o->lcic; ++k) win[size1 + k] *= o->etd_rk[ir];
  etd_h1rat_prod(o, ir, 1, 1, lf, win);
  for (size_t k = 0; k < sz; ++k) wout[k] = win[k] + wout[k];
}

/* cletdrk4::prop (:69-114), then the mask of tdse_prop12.  lf: the six
 * arrays of orc_etdrb_step (this scheme reads lfield(t), lfield(t + dt/2)
 * and lfield(t + dt); the rational functions take lfield(t + dt/2)). */
int orc_etdrk4_step(orc* o, const double* lf, zc* wfn, int apply_mask) {
  if (!o->etd_ready || o->etd_scheme != 2) return -1;
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  const double *lf0 = lf, *lf1 = lf + 9, *lf3 = lf + 18;
  zc* buf = xcalloc(sz * 9, sizeof(zc));
  zc *w0 = buf, *w1 = buf + sz, *w2 = buf + 2 * sz, *w3 = buf + 3 * sz, *hw0 = buf + 4 * sz, *hw1 = buf + 5 * sz,
     *hw2 = buf + 6 * sz, *hw3 = buf + 7 * sz, *tw = buf + 8 * sz;
  memcpy(w0, wfn, sizeof(zc) * sz);
  orc_htot(o, 1.0, lf0, w0, w0 + size1, hw0, hw0 + size1);
  etk_prod1(o, ETK_GFUN, lf1, hw0, w1);
  etk_prod1(o, ETK_EXP2, lf1, w0, tw);
  for (size_t k = 0; k < sz; ++k) w1[k] = tw[k] + w1[k];
  orc_htot(o, 1.0, lf1, w1, w1 + size1, hw1, hw1 + size1);
  etk_prod1(o, ETK_GFUN, lf1, hw1, w2);
  for (size_t k = 0; k < sz; ++k) w2[k] = tw[k] + w2[k];
  orc_htot(o, 1.0, lf1, w2, w2 + size1, hw2, hw2 + size1);
  for (size_t k = 0; k < sz; ++k) tw[k] = 2.0 * hw2[k] + (-1.0) * hw0[k];
  etk_prod1(o, ETK_GFUN, lf1, tw, w3);
  etk_prod1(o, ETK_EXP2, lf1, w1, tw);
  for (size_t k = 0; k < sz; ++k) w3[k] = tw[k] + w3[k];
  orc_htot(o, 1.0, lf3, w3, w3 + size1, hw3, hw3 + size1);
  for (size_t k = 0; k < sz; ++k) hw2[k] = hw1[k] + hw2[k];
  etk_prod1(o, ETK_EXP1, lf1, w0, wfn);
  etk_prod2(o, ETK_F0, lf1, hw0, wfn);
  etk_prod2(o, ETK_F1, lf1, hw2, wfn);
  etk_prod2(o, ETK_F2, lf1, hw3, wfn);
  if (apply_mask) orc_mask(o, wfn);
  free(buf);
  return 0;
}

/* ------------------------------------------------------------------ */
/* cllawrk (src/prop/cllawrk.cpp): Lawson RK4 (prop4, :188-280) with    */
/* the (1,1) Pade of cldpade for exp(-i h1 dt/2) (lawrk_npade 11,       */
/* h1itr_midpt true: expwfn11, :296-324), the defaults of gen (:20-75). */
/* ------------------------------------------------------------------ */
/* cllawrk::gen: dPade2p.gen(dt/2, alpha 2) = dpade_gen's band LU of
 * 1 + (i dt/4) T_l, the factor orc_h1itr_gen builds for its argument dt/2. */
int orc_lawrk_gen(orc* o, double dtime, int maxcyc, double thresh) {
  if (o->isplit != 1) return -2;
  int rc = orc_h1itr_gen(o, 0.5 * dtime);
  if (rc) return rc;
  o->law_dt = dtime;
  o->law_maxcyc = maxcyc;
  o->law_thresh = thresh;
  o->law_ready = 1;
  return 0;
}

/* dpade_prod (src/prop/dpade_prod.F90) with dtime dt/2, alpha 2, split_type
 * 1, iprojfc 2: x0 = (1 + i T dt/4)^-1 w, x_k = (1 + i T dt/4)^-1 (-i V dt/4)
 * x_(k-1) while the RMS orbital norm of x_(k-1) >= thresh (at most maxcyc
 * more terms), w(dyn) = sum x_k. */
static void law_dpade(orc* o, const double* lf, zc* w) {
  int nf = o->nfun, nfc = o->nfcore, nb = o->nbas;
  size_t size1 = (size_t)nb * nf;
  zc zfac = (o->icomp == 1 ? -I : (zc)-1.0) * (0.5 * o->law_dt) / 2.0 * lf[2];
  zc* orb = xcalloc(size1, sizeof(zc));
  zc* horb = xcalloc(size1, sizeof(zc));
  zc* torb = xcalloc(size1, sizeof(zc));
  memcpy(orb, w, size1 * sizeof(zc));
  h1itr_cninv(o, orb);
  memcpy(torb, orb, size1 * sizeof(zc));
  for (int icyc = 1; icyc <= o->law_maxcyc; ++icyc) {
    double rms = 0.0;
    for (int ifun = nfc; ifun < nf; ++ifun) {
      const zc* x = orb + (size_t)ifun * nb;
      for (int k = 0; k < nb; ++k) rms += creal(x[k]) * creal(x[k]) + cimag(x[k]) * cimag(x[k]);
    }
    rms = sqrt(rms / nf);
    if (rms < o->law_thresh) break;
    memset(horb, 0, size1 * sizeof(zc));
    if (o->igauge == 0)
      zprod(o, zfac, orb, horb, nfc, nf, o->nr);
    else
      pzprod(o, zfac, orb, horb, nfc, nf, o->nr);
    memcpy(orb, horb, size1 * sizeof(zc));
    h1itr_cninv(o, orb);
    for (size_t k = 0; k < size1; ++k) torb[k] += orb[k];
  }
  memcpy(w + (size_t)nb * nfc, torb + (size_t)nb * nfc, sizeof(zc) * (size1 - (size_t)nb * nfc));
  free(orb);
  free(horb);
  free(torb);
}

/* cllawrk::expwfn11: out = in, the Pade factor on out's orbitals with the
 * field at t + dt/2, then out.orb = -in.orb + 2 out.orb (the Cayley form of
 * exp(-i h1 dt/2)); the CI part is copied. */
static void law_expwfn(orc* o, const double* lf_half, const zc* in, zc* out) {
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  memcpy(out, in, sizeof(zc) * sz);
  law_dpade(o, lf_half, out);
  for (size_t k = 0; k < size1; ++k) out[k] = -1.0 * in[k] + 2.0 * out[k];
}

/* cllawrk::prop4, then the mask of tdse_prop12; lf as for orc_etdrb_step
 * (this scheme reads lfield(t), lfield(t + dt/2), lfield(t + dt)). */
int orc_lawrk_step(orc* o, const double* lf, zc* wfn, int apply_mask) {
  if (!o->law_ready) return -1;
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  const double *lf0 = lf, *lf1 = lf + 9, *lf3 = lf + 18;
  double dt = o->law_dt;
  zc cdt1 = dt, cdt2 = dt / 2.0, cdt3 = dt / 3.0, cdt6 = dt / 6.0;
  zc* buf = xcalloc(sz * 4, sizeof(zc));
  zc *w0 = buf, *w1 = buf + sz, *hw0 = buf + 2 * sz, *hw1 = buf + 3 * sz;
  memcpy(w0, wfn, sizeof(zc) * sz);
  orc_htot(o, 1.0, lf0, w0, w0 + size1, hw0, hw0 + size1);
  law_expwfn(o, lf1, w0, w1);
  law_expwfn(o, lf1, hw0, hw1);
  for (size_t k = 0; k < sz; ++k) wfn[k] = cdt6 * hw1[k] + w1[k];
  for (size_t k = 0; k < sz; ++k) w0[k] = cdt2 * hw1[k] + w1[k];
  orc_htot(o, 1.0, lf1, w0, w0 + size1, hw0, hw0 + size1);
  for (size_t k = 0; k < sz; ++k) wfn[k] += cdt3 * hw0[k];
  for (size_t k = 0; k < sz; ++k) w0[k] = cdt2 * hw0[k] + w1[k];
  orc_htot(o, 1.0, lf1, w0, w0 + size1, hw0, hw0 + size1);
  for (size_t k = 0; k < sz; ++k) wfn[k] += cdt3 * hw0[k];
  for (size_t k = 0; k < sz; ++k) w1[k] += cdt1 * hw0[k];
  law_expwfn(o, lf1, w1, w0);
  orc_htot(o, 1.0, lf3, w0, w0 + size1, hw0, hw0 + size1);
  law_expwfn(o, lf1, wfn, w1);
  for (size_t k = 0; k < sz; ++k) wfn[k] = cdt6 * hw0[k] + w1[k];
  if (apply_mask) orc_mask(o, wfn);
  free(buf);
  return 0;
}

/* ------------------------------------------------------------------ */
/* clexpab (src/prop/clexpab.cpp): the exponential Adams-Bashforth      */
/* method of prop_general (:370-496) with gen's defaults (:25-110):     */
/* expab_order 4, expab_ndpade 12 (the [1/2] form of gen_phi4,           */
/* :1148-1354), expab_pc false, expab_h1type -1.                         */
/* ------------------------------------------------------------------ */
#define XAB_K 4
#define XAB_RAT 7 /* its rational function's slot */

/* A^-1 of the n x n Taylor matrix A(i, j) = (-i)^j / j!, A(i, 0) = 1,
 * A(0, j >= 1) = 0 (column-major, as prop_general builds it), by
 * Gauss-Jordan with partial pivoting (the reference inverts it through
 * util_gmatinv_real, src/util/util_gmatinv.F90:46-87). */
static void xab_taylor_inverse(int n, double* ainv) {
  double a[XAB_K * XAB_K], b[XAB_K * XAB_K];
  for (int i = 0; i < n; ++i) a[i] = 1.0;
  for (int j = 1; j < n; ++j) a[n * j] = 0.0;
  double fact = 1.0;
  for (int j = 1; j < n; ++j) {
    fact *= j;
    for (int i = 1; i < n; ++i) a[n * j + i] = pow(-i, j) / fact;
  }
  for (int k = 0; k < n * n; ++k) b[k] = 0.0;
  for (int i = 0; i < n; ++i) b[n * i + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int pr = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(a[n * c + r]) > fabs(a[n * c + pr])) pr = r;
    for (int j = 0; j < n; ++j) {
      double t = a[n * j + c]; a[n * j + c] = a[n * j + pr]; a[n * j + pr] = t;
      t = b[n * j + c]; b[n * j + c] = b[n * j + pr]; b[n * j + pr] = t;
    }
    double piv = a[n * c + c];
    for (int j = 0; j < n; ++j) {
      a[n * j + c] /= piv;
      b[n * j + c] /= piv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[n * c + r];
      for (int j = 0; j < n; ++j) {
        a[n * j + r] -= f * a[n * j + c];
        b[n * j + r] -= f * b[n * j + c];
      }
    }
  }
  for (int k = 0; k < n * n; ++k) ainv[k] = b[k];
}

/* clexpab::gen: the [1/2] form of phi4 (dfac 6, numerator 1 - x, roots
 * 2 -+ sqrt(2) i), numerator scaled by dt, denominator slopes by tunit dt;
 * the history restarts. */
int orc_expab_gen(orc* o, double dtime, int maxcyc, double thresh) {
  if (o->isplit != 1) return -2;
  o->etd_maxcyc = maxcyc;
  o->etd_thresh = thresh;
  const double nn[2] = {1.0, -1.0};
  const zc roots[2] = {2.0 - 1.4142135623730950488 * I, 2.0 + 1.4142135623730950488 * I};
  if (etd_rat_set(o, XAB_RAT, 6.0, nn, 1, roots, 2, dtime, dtime)) return -1;
  xab_taylor_inverse(XAB_K, o->xab_ainv);
  size_t sz = (size_t)o->nbas * o->nfun + o->lcic;
  free(o->xab_buf);
  o->xab_buf = xcalloc(sz * (3 * XAB_K + 1), sizeof(zc));
  o->xab_nstep = 0;
  o->law_dt = dtime;
  o->xab_ready = 1;
  return 0;
}

/* clexpab::prop_general, then the mask of tdse_prop12.  lf: lfield(t) then
 * lfield(timeP[is]) for is = 1 .. K-1, the times of the history after this
 * step's shift (t itself while the history is being filled: the first step
 * copies the start state into every slot). */
int orc_expab_step(orc* o, const double* lf, zc* wfn, int apply_mask) {
  if (!o->xab_ready) return -1;
  const int K = XAB_K, K1 = XAB_K - 1;
  size_t size1 = (size_t)o->nbas * o->nfun, sz = size1 + o->lcic;
  zc tunit = o->icomp == 0 ? -1.0 : -I;
  double dt = o->law_dt;
  zc *P[XAB_K], *hP[XAB_K], *B[XAB_K], *tw = o->xab_buf + (size_t)3 * K * sz;
  for (int is = 0; is < K; ++is) {
    P[is] = o->xab_buf + (size_t)is * sz;
    hP[is] = o->xab_buf + (size_t)(K + is) * sz;
    B[is] = o->xab_buf + (size_t)(2 * K + is) * sz;
  }
  if (o->xab_nstep == 0) {
    memcpy(P[0], wfn, sizeof(zc) * sz);
    orc_htot(o, 1.0, lf, wfn, wfn + size1, hP[0], hP[0] + size1);
    for (int is = 1; is < K; ++is) {
      memcpy(P[is], P[0], sizeof(zc) * sz);
      memcpy(hP[is], hP[0], sizeof(zc) * sz);
    }
  } else {
    for (int is = K1; is > 0; --is) {
      memcpy(P[is], P[is - 1], sizeof(zc) * sz);
      memcpy(hP[is], hP[is - 1], sizeof(zc) * sz);
    }
    memcpy(P[0], wfn, sizeof(zc) * sz);
    orc_htot(o, 1.0, lf, wfn, wfn + size1, hP[0], hP[0] + size1);
  }
  o->xab_nstep++;
  for (int is = 0; is < K; ++is) memcpy(B[is], hP[is], sizeof(zc) * sz);
  for (int is = 1; is < K; ++is) {
    double lfp[9];
    for (int ii = 0; ii < 9; ++ii) lfp[ii] = lf[9 * is + ii] - lf[ii];
    etd_v1ext(o, tunit, lfp, P[is], B[is]);
  }
  etd_h1tot(o, lf, wfn, tw);
  for (size_t k = 0; k < sz; ++k) tw[k] = tunit * tw[k] + hP[0][k];
  memcpy(hP[K1], tw, sizeof(zc) * sz);
  int jfact = 1;
  for (int js = 1; js < K; ++js) {
    jfact *= js;
    zc rj = dt / jfact;
    for (size_t k = 0; k < sz; ++k) wfn[k] += rj * tw[k];
    memcpy(P[K1], tw, sizeof(zc) * sz);
    etd_h1tot(o, lf, P[K1], tw);
    for (size_t k = 0; k < size1; ++k) tw[k] *= tunit * dt;
    for (int is = 0; is < K; ++is) {
      zc af = o->xab_ainv[K * is + js];
      for (size_t k = 0; k < sz; ++k) tw[k] += af * B[is][k];
    }
  }
  jfact *= K;
  zc rj = dt / jfact;
  for (int k = 0; k < o->lcic; ++k) tw[size1 + k] *= rj;
  etd_h1rat_prod(o, XAB_RAT, 1, 1, lf, tw);
  for (size_t k = 0; k < sz; ++k) wfn[k] += tw[k];
  if (apply_mask) orc_mask(o, wfn);
  return 0;
}

/* ------------------------------------------------------------------ */
/* hprod_chkconv (src/hprod/hprod_chkconv.F90:2-63) and init_prop12's   */
/* convergence test (src/main/init.cpp:242-310)                         */
/* ------------------------------------------------------------------ */
/* out[0] the energy of an hprod (ene_tot); out[1] |hcic| / lcic; through
 * hprod_gfock, hprod_projall with inner false (src/hprod/hprod_projall.F90:
 * 2-160: fmat(i, j) = <phi_i|g_j> over r < nradfc, g_i -= sum_j phi_j
 * fmat(j, i) for the dynamical i) and hprod_mkovlp over the whole box
 * (ovlp(a, b) = <g_a|g_b>): out[2] = sqrt(sum_dyn ovlp(i, i)) / nfun,
 * out[3] = sqrt(2 sum_dcore ovlp(i, i) + sum_act ovlp(a, b) den1(b, a)). */
static void gfock(orc* o, const double* lfield, const zc* wfn, const zc* cic);

void orc_chkconv(orc* o, const double* lfield, const zc* wfn, const zc* cic, double* out) {
  int nf = o->nfun, nfc = o->nfcore, ncore = o->ncore, na = o->nact, nr = o->nr, L1 = o->L1;
  size_t nb = o->nbas, size1 = nb * nf;
  zc* hw = xcalloc(size1, sizeof(zc));
  zc* hc = xcalloc(o->lcic > 0 ? o->lcic : 1, sizeof(zc));
  orc_htot(o, 1.0, lfield, wfn, cic, hw, hc);
  out[0] = o->ene_tot;
  double hn = 0.0;
  for (int k = 0; k < o->lcic; ++k) hn += creal(conj(hc[k]) * hc[k]);
  out[1] = sqrt(hn) / o->lcic;
  gfock(o, lfield, wfn, cic);
  zc* fmat = xcalloc((size_t)nf * nf, sizeof(zc));
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < nf; ++j)
      if (o->mval[j] == o->mval[i]) {
        zc t = 0;
        for (int l = 0; l <= o->lmax1; ++l)
          for (int r = 0; r < o->nradfc; ++r) t += conj(o->orb[OIDX(r, l, i)]) * o->gorb[OIDX(r, l, j)];
        fmat[i + (size_t)nf * j] = t;
      }
  for (int i = nfc; i < nf; ++i)
    for (int j = 0; j < nf; ++j)
      if (o->mval[j] == o->mval[i])
        for (int l = 0; l <= o->lmax1; ++l)
          for (int r = 0; r < o->nradfc; ++r) o->gorb[OIDX(r, l, i)] -= o->orb[OIDX(r, l, j)] * fmat[j + (size_t)nf * i];
  zc* ov = xcalloc((size_t)nf * nf, sizeof(zc));
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < nf; ++j)
      if (o->mval[j] == o->mval[i]) {
        int am = abs(o->mval[i]);
        zc t = 0;
        for (int l = am; l < L1; ++l)
          for (int r = 0; r < nr; ++r) t += conj(o->gorb[OIDX(r, l, j)]) * o->gorb[OIDX(r, l, i)];
        ov[j + (size_t)nf * i] = t;
      }
  zc c2 = 0, c3 = 0;
  for (int i = nfc; i < nf; ++i) c2 += ov[i + (size_t)nf * i];
  for (int i = nfc; i < ncore; ++i) c3 += 2.0 * ov[i + (size_t)nf * i];
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja < na; ++ja) c3 += ov[(ncore + ia) + (size_t)nf * (ncore + ja)] * o->den1[ja + na * ia];
  out[2] = sqrt(creal(c2)) / nf;
  out[3] = sqrt(creal(c3));
  free(hw);
  free(hc);
  free(fmat);
  free(ov);
}

/* ------------------------------------------------------------------ */
/* accessors for tests                                                 */
/* ------------------------------------------------------------------ */
void orc_sizes(const orc* o, int* out) {
  out[0] = o->nrad;
  out[1] = o->nbas;
  out[2] = o->nfun;
  out[3] = o->lcic;
  out[4] = o->nlat;
  out[5] = o->L1;
  out[6] = o->L2;
  out[7] = o->nr;
  out[8] = o->mmax1;
  out[9] = o->mmax2;
}

const void* orc_table(const orc* o, const char* name) {
#define T(x) if (!strcmp(name, #x)) return o->x
  T(xrad); T(wrad); T(mask); T(radk); T(radk0); T(radp); T(cost); T(wlat);
  T(legf1); T(legb1); T(legf2); T(legb2); T(alph); T(tmat); T(pmat); T(d2ll);
  T(zfac); T(pzfac1); T(pzfac2); T(d2fac1); T(d2fac2); T(d2invr); T(d2rpl0); T(d2rpl1);
  T(orb); T(h0orb); T(h1orb); T(gorb); T(v2orb); T(orbg); T(gorbg); T(v2orbg);
  T(rho2); T(v2sph); T(v2ang); T(orb0); T(rhofc); T(v2jfc); T(den1); T(den2);
  T(rden); T(rrden); T(int1e); T(int2e); T(xmat); T(dcic); T(dvr_x); T(dvr_w); T(dshape);
#undef T
  return NULL;
}

/* Length-form dipole <z> (hprod_dipole dip(1), src/hprod/hprod_dipole.F90:
 * 2-146): orbitals as hprod_orbin builds them for lfield (frozen cores with
 * the velocity-gauge phase), den1 from the CI, z on every orbital
 * (hprod_zprod_all with zfac = 1) and hprod_trace1 with the frozen cores
 * included (src/hprod/hprod_trace1.F90:2-88). */
/* hprod_trace1 (src/hprod/hprod_trace1.F90:2-88), dofc = true, ecs off:
 * 2 sum_core <i|p_i> (frozen cores over r < nradfc) + sum_ab conj(D(b,a)) <b|p_a>
 * on the orbitals of the last orbin (o->orb) */
static zc trace1(const orc* o, const zc* pw) {
  int nb = o->nbas, nfc = o->nfcore, ncore = o->ncore, na = o->nact, nr = o->nr;
  zc t = 0;
  for (int ifun = 0; ifun < ncore; ++ifun) {
    int dim = ifun < nfc ? o->nradfc : nr;
    for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
      for (int i = 0; i < dim; ++i) t += conj(o->orb[OIDX(i, l, ifun)]) * pw[OIDX(i, l, ifun)];
  }
  t *= 2.0;
  zc* tw = xcalloc((size_t)nb, sizeof(zc));
  for (int ia = 0; ia < na; ++ia) {
    int ifun = ncore + ia;
    memset(tw, 0, sizeof(zc) * nb);
    for (int ja = 0; ja < na; ++ja) {
      int jfun = ncore + ja;
      if (o->mval[ifun] != o->mval[jfun]) continue;
      for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
        for (int i = 0; i < nr; ++i) tw[(size_t)l * nr + i] += o->orb[OIDX(i, l, jfun)] * o->den1[ja + na * ia];
    }
    for (int l = 0; l <= o->lmax1; ++l)
      for (int i = 0; i < nr; ++i) t += conj(tw[(size_t)l * nr + i]) * pw[OIDX(i, l, ifun)];
  }
  free(tw);
  return t;
}

/* hprod_azprodp (src/hprod/hprod_azprod.F90:106-136) with bas_azfac of a
 * bare nucleus (src/bas/bas_gen_azfac.F90, the plain branch):
 * azfac(r,l,m) = -alph(l,m) Z / r^2 */
static void azprod(orc* o, zc zf, const zc* wfn, zc* hwfn, int f0, int f1, int dim) {
  for (int ifun = f0; ifun < f1; ++ifun) {
    int m = o->mval[ifun];
    for (int l = abs(m); l <= o->lmax1 - 1; ++l) {
      double az = -ALPH(l, m) * o->znuc;
      for (int i = 0; i < dim; ++i) {
        double r = o->xrad[i + 1];
        zc tmp = zf * (az / (r * r));
        hwfn[OIDX(i, l, ifun)] += wfn[OIDX(i, l + 1, ifun)] * tmp;
        hwfn[OIDX(i, l + 1, ifun)] += wfn[OIDX(i, l, ifun)] * tmp;
      }
    }
  }
}

double orc_dipole(orc* o, const double* lfield, const zc* wfn, const zc* cic) {
  int nf = o->nfun, nb = o->nbas, nfc = o->nfcore, nr = o->nr;
  orc_htotx(o, 1.0, lfield, wfn, cic); /* orbin + den1 */
  zc* zo = xcalloc((size_t)nb * nf, sizeof(zc));
  if (nfc > 0) zprod(o, 1.0, o->orb, zo, 0, nfc, o->nradfc);
  zprod(o, 1.0, o->orb, zo, nfc, nf, nr);
  zc t = trace1(o, zo);
  free(zo);
  return creal(t);
}

/* hprod_dipole's Ehrenfest terms (src/hprod/hprod_dipole.F90:51-72):
 * out[0] = dip(1) = Re tr(D z), out[1] = vel(1) = Re tr(D p_z) (+ A_z N in
 * the velocity gauge), out[2] = acc(1) = Re tr(D a_z) - E_z N, with the
 * property vectors of hprod_zprod_all / hprod_pzprod_all / hprod_azprod_all
 * at zfac = 1 (frozen cores over r < nradfc) and N = neltot(3).  The
 * frozen-core correction terms dip(2), vel(2), acc(2), which need the
 * generalised Fock vector of hprod_gfock, are in orc_dipole_all. */
void orc_dipole3(orc* o, const double* lfield, const zc* wfn, const zc* cic, double* out) {
  int nf = o->nfun, nb = o->nbas, nfc = o->nfcore, nr = o->nr;
  orc_htotx(o, 1.0, lfield, wfn, cic); /* orbin + den1 */
  zc* zo = xcalloc((size_t)nb * nf, sizeof(zc));
  zc* po = xcalloc((size_t)nb * nf, sizeof(zc));
  zc* ao = xcalloc((size_t)nb * nf, sizeof(zc));
  if (nfc > 0) {
    zprod(o, 1.0, o->orb, zo, 0, nfc, o->nradfc);
    pzprod(o, 1.0, o->orb, po, 0, nfc, o->nradfc);
    azprod(o, 1.0, o->orb, ao, 0, nfc, o->nradfc);
  }
  zprod(o, 1.0, o->orb, zo, nfc, nf, nr);
  pzprod(o, 1.0, o->orb, po, nfc, nf, nr);
  azprod(o, 1.0, o->orb, ao, nfc, nf, nr);
  double ne = (double)o->nel[2];
  out[0] = creal(trace1(o, zo));
  out[1] = creal(trace1(o, po));
  if (o->igauge == 1) out[1] += lfield[8] * ne;
  out[2] = creal(trace1(o, ao)) - lfield[5] * ne;
  free(ao);
  free(po);
  free(zo);
}

/* hprod_trace1 (src/hprod/hprod_trace1.F90:2-88) with dofc: 2 sum over the
 * cores of <a_i|b_i> (frozen cores over r < nradfc, l >= |m|) plus, for each
 * active a, <sum_j a_j den1(j, a) | b_a> (m_j = m_a) */
static zc trace1ab(const orc* o, const zc* a, const zc* b) {
  int nb = o->nbas, nfc = o->nfcore, ncore = o->ncore, na = o->nact, nr = o->nr;
  zc t = 0;
  for (int ifun = 0; ifun < ncore; ++ifun) {
    int dim = ifun < nfc ? o->nradfc : nr;
    for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
      for (int i = 0; i < dim; ++i) t += conj(a[OIDX(i, l, ifun)]) * b[OIDX(i, l, ifun)];
  }
  t *= 2.0;
  zc* tw = xcalloc((size_t)nb, sizeof(zc));
  for (int ia = 0; ia < na; ++ia) {
    int ifun = ncore + ia;
    memset(tw, 0, sizeof(zc) * nb);
    for (int ja = 0; ja < na; ++ja) {
      int jfun = ncore + ja;
      if (o->mval[ifun] != o->mval[jfun]) continue;
      for (int l = abs(o->mval[ifun]); l <= o->lmax1; ++l)
        for (int i = 0; i < nr; ++i) tw[(size_t)l * nr + i] += a[OIDX(i, l, jfun)] * o->den1[ja + na * ia];
    }
    for (int l = 0; l <= o->lmax1; ++l)
      for (int i = 0; i < nr; ++i) t += conj(tw[(size_t)l * nr + i]) * b[OIDX(i, l, ifun)];
  }
  free(tw);
  return t;
}

/* hprod_gfock (src/hprod/hprod_gfock.F90:2-92; jfc_implicit, xfc_implicit)
 * and hprod_gfock_sum (:198-230): o->gorb <- (h0 + h1 + g + v2) phi of
 * every orbital, the frozen cores over r < nradfc, with the mean field
 * acting on the frozen cores as hprod_mfprod2_gfock's onFC branches
 * (src/hprod/hprod_mfprod2.F90: fcx1 :158-200, dcj_onFC :282-304,
 * dcx_onFC :343-365, actj_onFC :414-436, actx_onFC :489-513). */
static void gfock(orc* o, const double* lfield, const zc* wfn, const zc* cic) {
  int nf = o->nfun, nb = o->nbas, nfc = o->nfcore, nfc2 = o->nfcore2, ncore = o->ncore, na = o->nact;
  int nr = o->nr, nlat = o->nlat, ng = o->ngrid, dimf = o->nradfc;
  /* orbin, RDMs, the pair potentials and the mean field of the dynamical
   * orbitals as in the hprod */
  orc_htotx(o, 1.0, lfield, wfn, cic);
  /* 1e operators of hprod_gfock: tprod_all, zprod_all / pzprod_all (the
   * frozen cores over r < nradfc) */
  zc zfield = lfield[2];
  memset(o->h0orb, 0, sizeof(zc) * (size_t)nb * nf);
  memset(o->h1orb, 0, sizeof(zc) * (size_t)nb * nf);
  if (nfc > 0) tprod(o, o->orb, o->h0orb, 0, nfc, dimf);
  tprod(o, o->orb, o->h0orb, nfc, nf, nr);
  if (o->igauge == 0) {
    if (nfc > 0) zprod(o, zfield, o->orb, o->h1orb, 0, nfc, dimf);
    zprod(o, zfield, o->orb, o->h1orb, nfc, nf, nr);
  } else {
    if (nfc > 0) pzprod(o, zfield, o->orb, o->h1orb, 0, nfc, dimf);
    pzprod(o, zfield, o->orb, o->h1orb, nfc, nf, nr);
  }
  /* the mean field acting on the frozen cores (theta grid, r < nradfc) */
  if (nfc > 0 && o->nel[2] >= 2) {
    mkrho2(o, 2);
    poisson(o, 2);
    sph2ang2(o, 2);
    v2herm(o, 2);
    zc* v2 = o->v2ang;
#define V2G(ig, jfun, ifun) v2[((size_t)(ifun) * nf + (jfun)) * ng + (ig)]
#define WG(ig, ifun) o->orbg[(size_t)(ifun) * ng + (ig)]
    for (size_t k = 0; k < (size_t)ng * nfc; ++k) o->gorbg[k] = o->v2orbg[k] = 0;
    zc* d2v = xcalloc((size_t)ng * (na > 0 ? na : 1), sizeof(zc));
    for (int ilat = 0; ilat < nlat; ++ilat)
      for (int i = 0; i < dimf; ++i) {
        size_t ig = (size_t)ilat * nr + i;
        /* fcx1 on FC */
        if (o->nfcore1 > 0)
          for (int ifun = 0; ifun < nfc; ++ifun)
            for (int jfun = nfc2; jfun < nfc; ++jfun) o->gorbg[(size_t)ifun * ng + ig] -= V2G(ig, jfun, ifun) * WG(ig, jfun);
        if (o->ndcore > 0) {
          /* dcj on FC */
          zc dv = 0;
          for (int jfun = nfc; jfun < ncore; ++jfun) dv += V2G(ig, jfun, jfun) * 2.0;
          for (int ifun = 0; ifun < nfc; ++ifun) o->gorbg[(size_t)ifun * ng + ig] += dv * WG(ig, ifun);
          /* dcx on FC */
          for (int ifun = 0; ifun < nfc; ++ifun)
            for (int jfun = nfc; jfun < ncore; ++jfun) o->gorbg[(size_t)ifun * ng + ig] -= V2G(ig, jfun, ifun) * WG(ig, jfun);
        }
        if (ncore > 0 && na > 0) {
          /* actj on FC */
          zc dv = 0;
          for (int ja = 0; ja < na; ++ja)
            for (int ka = 0; ka < na; ++ka) dv += V2G(ig, ncore + ja, ncore + ka) * o->den1[ka + na * ja];
          for (int ifun = 0; ifun < nfc; ++ifun) o->v2orbg[(size_t)ifun * ng + ig] += dv * WG(ig, ifun);
          /* actx on FC */
          for (int ja = 0; ja < na; ++ja) {
            int jfun = ncore + ja;
            zc wd = 0;
            for (int ka = 0; ka < na; ++ka) {
              int kfun = ncore + ka;
              if (o->mval[jfun] == o->mval[kfun]) wd += WG(ig, kfun) * (-o->den1[ka + na * ja] * 0.5);
            }
            for (int ifun = 0; ifun < nfc; ++ifun) o->v2orbg[(size_t)ifun * ng + ig] += wd * V2G(ig, jfun, ifun);
          }
        }
      }
    free(d2v);
#undef V2G
#undef WG
    ang2sph1(o, o->gorbg, o->gorb, 0, nfc, dimf);
    ang2sph1(o, o->v2orbg, o->v2orb, 0, nfc, dimf);
  }
  /* hprod_gfock_sum */
  for (int ifun = 0; ifun < nf; ++ifun) {
    int dim = ifun < nfc ? dimf : nr;
    for (int l = 0; l <= o->lmax1; ++l)
      for (int i = 0; i < dim; ++i) {
        size_t q = OIDX(i, l, ifun);
        o->gorb[q] = o->v2orb[q] + o->gorb[q] + o->h1orb[q] + o->h0orb[q];
      }
  }
}

/* hprod_dipole (src/hprod/hprod_dipole.F90:2-146) with frozen cores: out =
 * {dip(1..3), vel(1..3), acc(1..3)}.  dip/vel/acc(1) are the Ehrenfest
 * traces (orc_dipole3); with frozen cores the generalised Fock vector f of
 * gfock gives the corrections (:82-125):
 *   vel(2) = -4 Im sum_{fc i} <z i|f i> - 2 Im tr(D z, P_fc f),
 *   acc(2) = -4 Im sum_{fc i} <p_z i|f i> - 2 Im tr(D p_z, P_fc f),
 * f += E_z z phi first in the velocity gauge (:60-63), P_fc the projector on
 * the frozen cores over r < nradfc (hprod_projfc, src/hprod/hprod_projfc.F90:2-158
 * with inner); dip(2) = 0 and (3) = (1) + (2). */
void orc_dipole_all(orc* o, const double* lfield, const zc* wfn, const zc* cic, double* out) {
  int nf = o->nfun, nb = o->nbas, nfc = o->nfcore, nr = o->nr, dimf = o->nradfc;
  if (nfc > 0) gfock(o, lfield, wfn, cic);
  else orc_htotx(o, 1.0, lfield, wfn, cic);
  zc* zo = xcalloc((size_t)nb * nf, sizeof(zc));
  zc* po = xcalloc((size_t)nb * nf, sizeof(zc));
  zc* ao = xcalloc((size_t)nb * nf, sizeof(zc));
  if (nfc > 0) {
    zprod(o, 1.0, o->orb, zo, 0, nfc, dimf);
    pzprod(o, 1.0, o->orb, po, 0, nfc, dimf);
    azprod(o, 1.0, o->orb, ao, 0, nfc, dimf);
  }
  zprod(o, 1.0, o->orb, zo, nfc, nf, nr);
  pzprod(o, 1.0, o->orb, po, nfc, nf, nr);
  azprod(o, 1.0, o->orb, ao, nfc, nf, nr);
  double ne = (double)o->nel[2];
  double dip1 = creal(trace1ab(o, o->orb, zo));
  double vel1 = creal(trace1ab(o, o->orb, po));
  if (o->igauge == 1) vel1 += lfield[8] * ne;
  double acc1 = creal(trace1ab(o, o->orb, ao)) - lfield[5] * ne;
  double vc1 = 0, ac1 = 0, vc2 = 0, ac2 = 0;
  if (nfc > 0) {
    zc* f = o->gorb;
    if (o->igauge == 1)
      for (size_t k = 0; k < (size_t)nb * nf; ++k) f[k] += lfield[5] * zo[k];
    /* first correction term */
    zc vp = 0, ap = 0;
    for (int ifun = 0; ifun < nfc; ++ifun)
      for (int l = 0; l <= o->lmax1; ++l)
        for (int i = 0; i < dimf; ++i) {
          size_t q = OIDX(i, l, ifun);
          vp += conj(zo[q]) * f[q];
          ap += conj(po[q]) * f[q];
        }
    vc1 = -4.0 * cimag(vp);
    ac1 = -4.0 * cimag(ap);
    /* second correction term: f <- P_fc f */
    zc* fm = xcalloc((size_t)nf * nf, sizeof(zc));
    for (int ifun = 0; ifun < nfc; ++ifun)
      for (int jfun = 0; jfun < nf; ++jfun) {
        if (o->mval[jfun] != o->mval[ifun]) continue;
        zc t = 0;
        for (int l = 0; l <= o->lmax1; ++l)
          for (int i = 0; i < dimf; ++i) t += conj(o->orb[OIDX(i, l, ifun)]) * f[OIDX(i, l, jfun)];
        fm[ifun + nf * jfun] = t;
      }
    memset(f, 0, sizeof(zc) * (size_t)nb * nf);
    for (int ifun = 0; ifun < nf; ++ifun)
      for (int jfun = 0; jfun < nfc; ++jfun) {
        if (o->mval[jfun] != o->mval[ifun]) continue;
        for (int l = 0; l <= o->lmax1; ++l)
          for (int i = 0; i < dimf; ++i) f[OIDX(i, l, ifun)] += o->orb[OIDX(i, l, jfun)] * fm[jfun + nf * ifun];
      }
    free(fm);
    vc2 = -2.0 * cimag(trace1ab(o, zo, f));
    ac2 = -2.0 * cimag(trace1ab(o, po, f));
  }
  out[0] = dip1; out[1] = 0.0; out[2] = dip1;
  out[3] = vel1; out[4] = vc1 + vc2; out[5] = vel1 + vc1 + vc2;
  out[6] = acc1; out[7] = ac1 + ac2; out[8] = acc1 + ac1 + ac2;
  free(ao);
  free(po);
  free(zo);
}

/* hprod_mkovlp (src/hprod/hprod_mkovlp.F90:2-91) on o->orb:
 * ovlp[j + nf*i] = sum_{l >= |m|} sum_{i_r < cut} conj(orb_j) orb_i for m_i == m_j,
 * cut = rows up to and including the first one with r > rmax (all if none) */
static void mkovlp(const orc* o, double rmax, zc* ovlp) {
  int nf = o->nfun, nr = o->nr, cut = nr;
  for (int i = 0; i < nr; ++i)
    if (o->xrad[i + 1] > rmax) {
      cut = i + 1;
      break;
    }
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < nf; ++j) {
      zc t = 0;
      if (o->mval[i] == o->mval[j])
        for (int l = abs(o->mval[i]); l <= o->lmax1; ++l)
          for (int r = 0; r < cut; ++r) t += conj(o->orb[OIDX(r, l, j)]) * o->orb[OIDX(r, l, i)];
      ovlp[j + (size_t)nf * i] = t;
    }
}

/* determinant of an n x n complex matrix (row-major, destroyed): Gaussian
 * elimination with partial pivoting (util_det's role) */
static zc zdet(int n, zc* a) {
  zc d = 1;
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (cabs(a[i * n + k]) > cabs(a[p * n + k])) p = i;
    if (a[p * n + k] == 0) return 0;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        zc t = a[k * n + j];
        a[k * n + j] = a[p * n + j];
        a[p * n + j] = t;
      }
      d = -d;
    }
    d *= a[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      zc f = a[i * n + k] / a[k * n + k];
      for (int j = k; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
    }
  }
  return d;
}

/* One spin of ormas_dets0 / ormas_dett / ormas_dett_1
 * (src/ormas/ormas_dets0.F90:2-55; ormas_dett.F90:2-290; ormas_dett_1.F90:2-38):
 * electrons are the ncore cores then the active orbitals of the string in
 * ascending order; S[e][f] = ovlp(bra orbital e, ket orbital f).  d[k] is
 * the sum over k-subsets K of electrons of det(S with the rows of K replaced
 * by the orbital-identity rows delta(bra_e, ket_f)): the minors sums of
 * ormas_dett1/2/3 (k = 0: ormas_dets0, k = n: 1 if the strings are equal). */
static void spin_minors(const orc* o, const zc* ovlp, uint32_t si, uint32_t sj, int n, int kmax, zc* d) {
  int nf = o->nfun, nc = o->ncore;
  int bra[32], ket[32];
  int e = 0, f = 0;
  for (int c = 0; c < nc; ++c) bra[e++] = ket[f++] = c;
  for (int p = 0; p < o->nact; ++p) {
    if (si >> p & 1) bra[e++] = nc + p;
    if (sj >> p & 1) ket[f++] = nc + p;
  }
  for (int k = 0; k <= kmax; ++k) d[k] = 0;
  zc* m = xcalloc((size_t)n * n + 1, sizeof(zc));
  for (uint32_t K = 0; K < (1u << n); ++K) {
    int k = popcount32(K);
    if (k > kmax) continue;
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        m[a * n + b] = (K >> a & 1) ? (bra[a] == ket[b] ? 1.0 : 0.0) : ovlp[bra[a] + (size_t)nf * ket[b]];
    d[k] += n ? zdet(n, m) : 1.0;
  }
  free(m);
}

static double binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  double b = 1;
  for (int i = 1; i <= k; ++i) b = b * (n - k + i) / i;
  return b;
}

/* hprod_ipx (src/hprod/hprod_ipx.F90:2-19) -> ormas_ipx (src/ormas/ormas_ipx.F90)
 * -> ormas_ipx_core / ormas_ipx_ras (src/ormas/ormas_ipx_core.F90:2-130;
 * src/ormas/ormas_ipx_ras.F90:2-180): with the windowed overlap of the
 * orbitals of hprod_orbin, T(k) = sum_{ka+kb=k} Re sum_IJ conj(C_I) C_J
 * d^a_ka(I,J) d^b_kb(I,J) (T(k): k electrons free, the others inside), then
 * for k = max_ipx down to 1: ipx(k) = T(k) + sum_{j=1..k} (-1)^j C(N-k+j, j)
 * ipx(k-j), so that ipx(k) is the probability that exactly k electrons are
 * outside rad_ipx.  Returns -1 on a bad max_ipx. */
int orc_ipx(orc* o, int max_ipx, double rad_ipx, const double* lfield, const zc* wfn, const zc* cic, double* ipx) {
  int nf = o->nfun, nsa = o->nstr_a, nsb = o->nstr_b, N = o->nel[2];
  if (max_ipx < 0 || max_ipx > 8 || max_ipx > N) return -1;
  if (o->nact == 0) nsa = nsb = 1; /* cores only: one determinant, C = 1 (ormas_ipx_core) */
  orc_htotx(o, 1.0, lfield, wfn, cic); /* orbin */
  zc* ovlp = xcalloc((size_t)nf * nf, sizeof(zc));
  mkovlp(o, rad_ipx, ovlp);
  int na = o->nel[0], nb = o->nel[1];
  int ka_max = na < max_ipx ? na : max_ipx, kb_max = nb < max_ipx ? nb : max_ipx;
  zc* da = xcalloc((size_t)nsa * nsa * (ka_max + 1), sizeof(zc));
  zc* db = xcalloc((size_t)nsb * nsb * (kb_max + 1), sizeof(zc));
  for (int i = 0; i < nsa; ++i)
    for (int j = 0; j < nsa; ++j)
      spin_minors(o, ovlp, o->nact ? o->str_a[i] : 0, o->nact ? o->str_a[j] : 0, na, ka_max,
                  da + ((size_t)i * nsa + j) * (ka_max + 1));
  for (int i = 0; i < nsb; ++i)
    for (int j = 0; j < nsb; ++j)
      spin_minors(o, ovlp, o->nact ? o->str_b[i] : 0, o->nact ? o->str_b[j] : 0, nb, kb_max,
                  db + ((size_t)i * nsb + j) * (kb_max + 1));
  double T[9] = {0};
  for (int ib = 0; ib < nsb; ++ib)
    for (int ia = 0; ia < nsa; ++ia) {
      int dI = o->nact == 0 ? 0 : o->det_of[(size_t)ib * nsa + ia];
      if (dI < 0) continue;
      zc cI = o->nact == 0 ? 1.0 : cic[dI];
      for (int jb = 0; jb < nsb; ++jb)
        for (int ja = 0; ja < nsa; ++ja) {
          int dJ = o->nact == 0 ? 0 : o->det_of[(size_t)jb * nsa + ja];
          if (dJ < 0) continue;
          zc cJ = o->nact == 0 ? 1.0 : cic[dJ];
          const zc* pa = da + ((size_t)ia * nsa + ja) * (ka_max + 1);
          const zc* pb = db + ((size_t)ib * nsb + jb) * (kb_max + 1);
          for (int k = 0; k <= max_ipx; ++k) {
            zc s = 0;
            for (int ka = 0; ka <= ka_max; ++ka) {
              int kb = k - ka;
              if (kb >= 0 && kb <= kb_max) s += pa[ka] * pb[kb];
            }
            T[k] += creal(conj(cI) * cJ * s);
          }
        }
    }
  for (int k = 0; k <= max_ipx; ++k) ipx[k] = T[k];
  for (int k = max_ipx; k >= 1; --k)
    for (int j = 1; j <= k; ++j) ipx[k] += ipx[k - j] * binom(N - k + j, j) * ((j & 1) ? -1.0 : 1.0);
  free(db);
  free(da);
  free(ovlp);
  return 0;
}

/* hprod_ipd (src/hprod/hprod_ipx.F90:21-39) -> ormas_ipd
 * (src/ormas/ormas_ipd.F90:2-123): with the outer overlap
 * ovlx = 1 - ovlp (hermitised from the upper triangle), S1 and S2 from the
 * 1- and 2-RDM (cores doubly occupied), p0 = 1 - S1 + S2, p1 = S1 - 2 S2,
 * p2 = S2; ipd[k] for k = 0..max_ipd.  Returns -1 on a bad max_ipd. */
int orc_ipd(orc* o, int max_ipd, double rad_ipx, const double* lfield, const zc* wfn, const zc* cic, double* ipd) {
  int nf = o->nfun, nc = o->ncore, na = o->nact;
  if (max_ipd < 0 || max_ipd > 2 || max_ipd > o->nel[2]) return -1;
  orc_htotx(o, 1.0, lfield, wfn, cic); /* orbin + den1 + den2 */
  zc* ovlp = xcalloc((size_t)nf * nf, sizeof(zc));
  zc* ovlx = xcalloc((size_t)nf * nf, sizeof(zc));
  mkovlp(o, rad_ipx, ovlp);
#define OV(j, i) ovlx[(j) + (size_t)nf * (i)]
  for (int i = 0; i < nf; ++i) {
    for (int j = 0; j < i; ++j) {
      OV(j, i) = -ovlp[j + (size_t)nf * i];
      OV(i, j) = conj(OV(j, i));
    }
    OV(i, i) = 1.0 - ovlp[i + (size_t)nf * i];
  }
  zc s1c = 0, s1a = 0, s2cc = 0, s2ca = 0, s2aa = 0;
  for (int i = 0; i < nc; ++i) s1c += 2.0 * OV(i, i);
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja < na; ++ja) s1a += o->den1[ja + na * ia] * OV(nc + ia, nc + ja);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) s2cc += 2.0 * OV(i, i) * OV(j, j) - OV(i, j) * OV(j, i);
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja < na; ++ja) {
      int i = nc + ia, j = nc + ja;
      zc t = 0;
      for (int k = 0; k < nc; ++k) t += 2.0 * OV(i, j) * OV(k, k) - OV(i, k) * OV(k, j);
      s2ca += o->den1[ja + na * ia] * t;
    }
  for (int ia = 0; ia < na; ++ia)
    for (int ja = 0; ja < na; ++ja)
      for (int ka = 0; ka < na; ++ka)
        for (int la = 0; la < na; ++la)
          s2aa += o->den2[ka + na * (ia + na * (la + na * ja))] * OV(nc + ia, nc + ka) * OV(nc + ja, nc + la);
#undef OV
  s2aa *= 0.5;
  zc s1 = s1c + s1a, s2 = s2cc + s2ca + s2aa;
  double p[3] = {creal(1.0 - s1 + s2), creal(s1 - 2.0 * s2), creal(s2)};
  for (int k = 0; k <= max_ipd; ++k) ipd[k] = p[k];
  free(ovlx);
  free(ovlp);
  return 0;
}

/* clwfn::orth with orth_type "schmidt" (src/wfn/clwfn.cpp:313-321):
 * wfn_orth_cic (src/wfn/wfn_orth.F90:2-29) normalises the CI vector;
 * wfn_orth_orb (:31-176, fedvr_normalized, ecs off) orthonormalises the
 * dynamical orbitals in order by modified Gram-Schmidt against every earlier
 * orbital of the same m (frozen cores included), over l >= |m| and all rows. */
void orc_orth(orc* o, zc* orb, zc* cic) {
  if (o->lcic > 0 && cic) {
    double nrm = 0;
    for (int d = 0; d < o->lcic; ++d) nrm += creal(conj(cic[d]) * cic[d]);
    nrm = 1.0 / sqrt(nrm);
    for (int d = 0; d < o->lcic; ++d) cic[d] *= nrm;
  }
  int nr = o->nr;
  for (int i = o->nfcore; i < o->nfun; ++i) {
    int mi = o->mval[i];
    for (int j = 0; j < i; ++j) {
      if (o->mval[j] != mi) continue;
      zc t = 0;
      for (int l = abs(mi); l <= o->lmax1; ++l)
        for (int r = 0; r < nr; ++r) t += conj(orb[OIDX(r, l, j)]) * orb[OIDX(r, l, i)];
      for (int l = abs(mi); l <= o->lmax1; ++l)
        for (int r = 0; r < nr; ++r) orb[OIDX(r, l, i)] -= orb[OIDX(r, l, j)] * t;
    }
    double nrm = 0;
    for (int l = abs(mi); l <= o->lmax1; ++l)
      for (int r = 0; r < nr; ++r) nrm += creal(conj(orb[OIDX(r, l, i)]) * orb[OIDX(r, l, i)]);
    nrm = 1.0 / sqrt(nrm);
    for (int l = abs(mi); l <= o->lmax1; ++l)
      for (int r = 0; r < nr; ++r) orb[OIDX(r, l, i)] *= nrm;
  }
}

/* cldvr::get_val (src/dvr/cldvr.cpp:223-235): the Lagrange polynomial of node
 * m of element e at r, over the element's ndvr+1 nodes. */
static double dvr_val(const orc* o, int e, int m, double r) {
  const double* xp = o->dvr_x + (size_t)e * (o->ndvr + 1);
  double v = 1.0;
  for (int k = 0; k <= o->ndvr; ++k)
    if (k != m) v *= (r - xp[k]) / (xp[m] - xp[k]);
  return v;
}

/* clfedvr::get_ife (src/fedvr/clfedvr.cpp:628-638). */
static int dvr_ife(const orc* o, double r) {
  for (int i = 0; i < o->nfe; ++i)
    if (r + 1e-10 < o->x1[i]) return i;
  return -1;
}

/* clbas::proj (src/bas/clbas.cpp:280-332) from grid a to grid b, then the CI
 * copy of guess read_proj (src/main/guess.cpp:47).  Loops run over the orbitals
 * and l both problems hold (the reference runs them to the larger counts);
 * a target point beyond a's last edge (ife = -1, where the reference indexes
 * element -1) gets nothing.  CI: the first min(lcic) coefficients. */
void orc_wfn_proj(const orc* a, const zc* orb1, const zc* cic1, const orc* b, zc* orb2, zc* cic2) {
  int nfun = a->nfun < b->nfun ? a->nfun : b->nfun;
  int lmax = a->lmax1 < b->lmax1 ? a->lmax1 : b->lmax1;
  memset(orb2, 0, sizeof(zc) * (size_t)b->nfun * b->nbas);
  for (int irad2 = 1; irad2 < b->nrad; ++irad2) {
    double rval = b->xrad[irad2], swgt = sqrt(b->wrad[irad2]);
    int ife1 = dvr_ife(a, rval);
    if (ife1 < 0) continue;
    for (int m = (1 - ife1 > 0 ? 1 - ife1 : 0); m < a->ndvr; ++m) {
      int irad1 = a->mapf[ife1] + m;
      double mat = dvr_val(a, ife1, m, rval) * swgt;
      for (int ifun = 0; ifun < nfun; ++ifun)
        for (int l = 0; l <= lmax; ++l)
          orb2[(size_t)ifun * b->nbas + (size_t)l * b->nr + irad2 - 1] +=
              orb1[(size_t)ifun * a->nbas + (size_t)l * a->nr + irad1 - 1] * mat;
    }
    if (ife1 < a->nfe - 1) {
      int irad1 = a->mapf[ife1 + 1];
      double mat = dvr_val(a, ife1, a->ndvr, rval) * swgt;
      for (int ifun = 0; ifun < nfun; ++ifun)
        for (int l = 0; l <= lmax; ++l)
          orb2[(size_t)ifun * b->nbas + (size_t)l * b->nr + irad2 - 1] +=
              orb1[(size_t)ifun * a->nbas + (size_t)l * a->nr + irad1 - 1] * mat;
    }
  }
  if (cic2) {
    memset(cic2, 0, sizeof(zc) * b->lcic);
    int nci = a->lcic < b->lcic ? a->lcic : b->lcic;
    if (cic1) memcpy(cic2, cic1, sizeof(zc) * nci);
  }
}

void orc_energies(const orc* o, double* e) {
  e[0] = o->ene_fcore;
  e[1] = o->ene_dcore;
  e[2] = o->ene_core;
  e[3] = o->ene_act;
  e[4] = o->ene_tot;
}

/* band Cholesky helpers exposed for pin tests */
int orc_band_chol(int n, int kd, double* ab) { return band_chol(n, kd, ab); }
int orc_band_lu(int n, int kd, zc* ab, int* ipiv) { return band_lu(n, kd, ab, ipiv); }
void orc_band_lu_solve(int n, int kd, const zc* ab, const int* ipiv, zc* b) { band_lu_solve(n, kd, ab, ipiv, b); }
void orc_band_chol_solve(int n, int kd, const double* ab, double* x) { band_chol_solve(n, kd, ab, x); }
void orc_herm_eig(int n, const zc* a, double* w, zc* u) { herm_eig(n, a, w, u); }
void orc_ci_sigma(orc* o, const zc* cic, zc* sig, const zc* int1e, const zc* int2e) {
  int n = o->nact;
  memcpy(o->int1e, int1e, sizeof(zc) * n * n);
  memcpy(o->int2e, int2e, sizeof(zc) * n * n * n * n);
  memset(sig, 0, sizeof(zc) * o->lcic);
  ci_sigma(o, cic, sig);
}
void orc_ci_mkden(orc* o, const zc* cic) { ci_mkden(o, cic); }

/* CI layout: the occupation bit patterns of the ordered alpha and beta
 * strings (bit k = active orbital k), the string pair of every determinant
 * and the ormas_cic0 determinant; returns lcic (1 and no strings without
 * active orbitals). */
int orc_ci_layout(const orc* o, uint32_t* str_a, uint32_t* str_b, int* det_a, int* det_b, int* cic0) {
  if (o->nact == 0) {
    *cic0 = 0;
    return 1;
  }
  if (str_a) memcpy(str_a, o->str_a, o->nstr_a * sizeof(uint32_t));
  if (str_b) memcpy(str_b, o->str_b, o->nstr_b * sizeof(uint32_t));
  if (det_a) memcpy(det_a, o->det_a, o->lcic * sizeof(int));
  if (det_b) memcpy(det_b, o->det_b, o->lcic * sizeof(int));
  *cic0 = o->cic0;
  return o->lcic;
}
int orc_nstr(const orc* o, int spin) { return o->nact == 0 ? 0 : (spin == 0 ? o->nstr_a : o->nstr_b); }
