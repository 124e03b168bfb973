// Host-side set-up of one td1c problem: FEDVR grid, angular quadrature,
// operator tables in the layouts the gfx950 kernels read, the element Schur
// complement factors of the radial Poisson operator, and the CAS string
// tables of the CI island.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/td1c_hprod.h"

namespace td1c {

struct Setup {
  // problem
  int nfe = 0, n = 0, nrad = 0, nr = 0, nradfc = 0;
  int lmax1 = 0, lmax2 = 0, L1 = 0, L2 = 0, nlat = 0, mmax1 = 0, mmax2 = 0;
  double znuc = 0;
  int nfcore2 = 0, nfcore1 = 0, nfcore = 0, ndcore = 0, ncore = 0, nact = 0, nfun = 0, ndyn = 0;
  std::vector<int> mval, lval;
  int nel[3] = {0, 0, 0}, nelact[3] = {0, 0, 0}, mtot = 0;
  int icomp = 1, igauge = 0, isplit = 0, ioorot = 0, iprojfc = 2;
  int jfc_implicit = 1, xfc_implicit = 1;
  double throcc1 = 1e-15, throcc2 = 1e-15, rmask = 0, mask_order = 4;
  int precision = 0;
  int lconst = 0;  // 0 off, 1 every orbital (wfn_ladapt), 2 the cores (wfn_ladapt_core)

  // FEDVR grid (clfedvr): global points 0..nrad
  std::vector<double> x0, x1;
  std::vector<double> xrad, wrad, mask;  // nrad+1
  // element-local DVR data per element: nodes, weights, derivative table
  std::vector<double> ex, ew, eds;  // nfe*(n+1), nfe*(n+1), nfe*(n+1)^2
  // global kinetic / derivative bands, band[(2n+1)*row + n + col - row]
  std::vector<double> radk, radk0, radp;

  // kernel tables ---------------------------------------------------------
  std::vector<double> xr;        // interior r, nr
  std::vector<double> bandT_k;   // [2n+1][nr]: radk row band, transposed (l = 0)
  std::vector<double> bandT_k0;  // [2n+1][nr]: radk0 row band (l >= 1)
  std::vector<double> bandT_p;   // [2n+1][nr]: radp row band
  std::vector<double> tdiag;     // [L1][nr] complex (re,im): diagonal of tmat
  std::vector<double> tzor;      // [L1][nr] complex (re,im): tmat diagonal without the kinetic part
  // element-local matrices [u][k][m] (m fastest) of the distinct elements
  // (equal element widths give bitwise-equal blocks, stored once; element e
  // uses block emap[e]): kinetic for l = 0 (radk), kinetic for l >= 1
  // (radk0), d/dr (radp); rows/columns of r = 0 and r = R are zero
  std::vector<double> eK0, eK, eP;
  std::vector<int> emap;
  std::vector<double> alph;      // [2mmax1+1][L1]

  // angular quadrature and transform matrices, [m-index][K][C]
  std::vector<double> cost, wlat;
  std::vector<double> mat_l2g;   // K = L1, C = nlat  (legf1)
  std::vector<double> mat_g2l;   // K = nlat, C = L1  (legb1)
  std::vector<double> mat_g2L;   // K = nlat, C = L2  (legb2)
  std::vector<double> mat_L2g;   // K = L2, C = nlat  (legf2)

  // Poisson: per L scalings and element Schur factors
  std::vector<double> d2fac1;                   // [L2]
  std::vector<double> d2invr, d2rpl0, d2rpl1, d2fac2;  // [L2][nr]
  int ni = 0;                                   // interior points per element = n-1
  std::vector<double> pG;    // [L2][nfe][ni][ni]  inverse interior blocks
  std::vector<double> pCl;   // [L2][nfe][ni]      A(bridge_left, interior)
  std::vector<double> pCr;   // [L2][nfe][ni]      A(bridge_right, interior)
  std::vector<double> pWl;   // [L2][nr]           G*A(interior, bridge_left) scattered
  std::vector<double> pWr;   // [L2][nr]
  std::vector<double> pSd;   // [L2][nfe-1]  Thomas: modified diagonal (pivots)
  std::vector<double> pSl;   // [L2][nfe-1]  Thomas: multipliers l_k = S(k,k-1)/piv(k-1)
  std::vector<double> pSu;   // [L2][nfe-1]  S(k,k+1)
  // parallel cyclic reduction of the same Schur system, per L: step k
  // (stride 2^k) updates d_i += pa[i] d_(i-h) + pg[i] d_(i+h); after
  // pcr_steps steps x_i = d_i * pbinv[i]
  int pcr_steps = 0;
  std::vector<double> pcrA, pcrG;  // [L2][pcr_steps][nfe-1]
  std::vector<double> pcrBinv;     // [L2][nfe-1]

  // orbital pair list evaluated per hprod: (j, i) as in hprod_mkrho2p
  struct Pair { int j, i, type; };  // type 0 = dyn-dyn, 1 = dyn-fc1
  std::vector<Pair> pairs;
  std::vector<int> pair_of;   // [nfun*nfun]: index of stored pair for unordered (j,i), -1 none
  std::vector<int> pair_conj; // [nfun*nfun]: 1 if v(j,i) = conj(stored)

  // CAS determinant space (nsub = 1)
  int nstr_a = 1, nstr_b = 1, lcic = 1;
  std::vector<uint32_t> str_a, str_b;
  std::vector<int> det_of;   // [nstr_b*nstr_a] -> CI index or -1
  int cic0 = 0;              // the determinant ormas_cic0 sets to 1
  // single-excitation tables on the full string spaces:
  // exA[(ia*nact + p)*nact + q] = target string index, sgA = sign (0 if none)
  std::vector<int> exA, exB;
  std::vector<int> sgA, sgB;

  std::string error;
  bool build(const td1c_problem& p);
  bool build_ci();  // the CAS string tables (called by build)
};

// FEDVR nodes and weights of nfe elements of order n on edges[0..nfe]:
// ex / ew [nfe][n+1] per element, xrad / wrad [n*nfe+1] global points with the
// bridge weights summed.
void fedvr_nodes(int nfe, int n, const double* edges, std::vector<double>& ex, std::vector<double>& ew,
                 std::vector<double>& xrad, std::vector<double>& wrad);

// Crank-Nicolson factors of the one-body half step of split_type = 1
// (clh1itr, h1itr_gen_cninv: 1 + fac T_l per l), in the element Schur form
// of the Poisson solve: per (l, e) the inverse interior block G (complex
// symmetric), W_l / W_r = G C^T scattered over the interior rows, and the
// complex symmetric tridiagonal bridge system by parallel cyclic reduction
// (step k updates d_i += pa[i] d_(i-h) + pg[i] d_(i+h); x_i = d_i binv[i]).
// Complex values are (re, im) pairs of doubles.
struct CnTables {
  int L1 = 0, nfe = 0, ni = 0, nr = 0, nb = 0, K = 0;
  std::vector<double> G;      // [L1][nfe][ni][ni] complex
  std::vector<double> Wl, Wr; // [L1][nr] complex
  std::vector<double> pa, pg; // [L1][K][nb] complex
  std::vector<double> binv;   // [L1][nb] complex
};

// T_l = the kinetic band (radk for l = 0, radk0 above) + the diagonal
// tdiag + v2jfc (frozen-core Hartree potential, nr complex values, or null).
bool build_cn(const Setup& S, const double* v2jfc, double fac_re, double fac_im, CnTables& T,
              std::string& error);

}  // namespace td1c
