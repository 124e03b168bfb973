// Launch arguments shared between the host orchestration (api.hip) and the
// gfx950 kernels (kernels.hip, island.hip).  Every pointer is device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "field.hpp"

namespace td1c {

constexpr int kMaxFun = 16;       // orbitals handled by the theta-grid mean-field kernel
constexpr int kMaxAct = 8;        // active orbitals handled on the theta grid
constexpr int kMaxActPairs = kMaxAct * kMaxAct;
constexpr int kMaxStages = 6;
constexpr int kMaxBridge = 160;   // nfe - 1 bridges handled by the Poisson bridge kernel
constexpr int kMaxPcrSteps = 8;   // ceil(log2(kMaxBridge))

// Time of a device-driven propagation: step n runs at t0 + n * dt.
struct PropClock {
  double t0, dt;
  long long n;
};

// Runge-Kutta stage nodes c_s.
struct StageNodes {
  int ns;
  double c[kMaxStages];
};

struct TransformBatch {
  long long in0;   // element offset of the input block (MODE 1: orbital j)
  long long in1;   // MODE 1: orbital i
  long long out;   // element offset of the output block
  int mat;         // matrix index (m + mmax)
  int rows;        // radial rows to compute (nr, or nradfc for frozen core)
};

struct OneBodyArgs {
  int nr, L1, ndvr, nfcore, nradfc, f0, f1, mmax1, igauge, fc_full;
  const int* mval;
  const double* band_k;   // [2n+1][nr] kinetic band for l = 0
  const double* band_k0;  // [2n+1][nr] kinetic band for l >= 1
  const double* band_p;   // [2n+1][nr] d/dr band
  const double* xr;       // [nr]
  const double* alph;     // [2mmax1+1][L1]
  const double2* tdiag;   // [L1][nr]
  const double* lf;       // device lfield[9]: [2] gauge field, [5] E (frozen-core length term)
};

struct OneBodyElArgs {
  int nr, nfe, L1, ndvr, nfcore, nradfc, nfun, mmax1, igauge, fc_full;
  const int* mval;
  const double *eK0, *eK, *eP;  // [u][k][m] element matrices of the distinct elements
  const int* emap;              // [nfe] element -> its block u
  const double* xr;
  const double* alph;
  const double2* tzor;          // [L1][nr] tmat diagonal without the kinetic part (+ v2jfc)
  const double* lf;             // device lfield[9]
  // right-hand sides of the clh1itr half step (cn != 0): the block's orbital
  // blockIdx.z is dynamical orbital fshift + blockIdx.z of the problem;
  // cn = 1: out = psi + cv (T psi + [isplit = 1] V psi), cn = 2: out = cv V psi
  // (returns at once when stop[0] is set), cn = 3: out = cv (T psi + V psi)
  int cn, fshift, isplit;
  double2 cv;
  const int* stop;
  // cn = 2 with chk_part set: first the convergence test of the previous
  // term (the sum of its chk_n block partial norms in k_cn_check's order,
  // rms over chk_nfun orbitals below chk_thresh): when it passes, chk_flag
  // is set (by one block) and the launch returns
  const double* chk_part;
  int chk_n, chk_nfun;
  double chk_thresh;
  int* chk_flag;
};

// One half step of clh1itr (split_type = 1; h1itr_prop,
// src/prop/h1itr_prop.F90:2-134): x0 = (1 + c T)^-1 (1 - c (T + V)) psi,
// x_k = (1 + c T)^-1 (-c V) x_(k-1), psi <- sum x_k over the dynamical
// orbitals, c = i dt/4 (icomp 1) or dt/4, V the field coupling of lf[2].
// (1 + c T_l)^-1 is the element Schur solve of the complex tables below.
struct CnArgs {
  int nr, n, nfe, ni, nb, K, L1, nfun, nfcore, mmax1, igauge, isplit;
  const int* mval;
  const double *band_k, *band_k0, *band_p;  // [2n+1][nr] transposed bands
  const double2* tdiag;                     // [L1][nr] tmat diagonal (+ v2jfc)
  const double *xr, *alph;
  const double* lf;                         // device lfield[9] of this half step
  const double2 *G;                         // [L1][gstride][k][i] = G_e(i, k)
  int gstride;                              // elements per l in G (nfe, or more in a shared table)
  const double2 *Wl, *Wr;                   // [L1][nr]
  const double2 *pa, *pg;                   // [L1][K][nb]
  const double2* binv;                      // [L1][nb]
  double2* wfn;                             // orbitals of the state (dynamical rows updated)
  double2* x;                               // [nd][L1][nr] current term
  double2* y;                               // [nd][L1][nr] right-hand side
  double2 zfac1;                            // -c
  double2 zfac0;                            // mode 2: the first right-hand side zfac0 psi
  double thresh;
  int* flag;                                // [0] converged
  double* part;                             // norm partial sums: bridge jobs, then interior blocks
  int npart1;
};
constexpr int kCnGroup = 8;             // dynamical orbitals per interior block
constexpr int kCnInteriorThreads = 256;
// partial sums of the interior pass for a problem (for the buffer size)
int cn_interior_blocks(const CnArgs& a);
// mode 0: the first term from psi; mode 1: one more term of the series
// (every kernel returns at once once flag[0] is set); mode 2: the first term
// of a plain solve, right-hand side zfac0 psi; mode 3: only the right-hand
// side y = zfac1 (T + V) x, an application of h1 (x, y of the arguments);
// mode 4: only y = zfac1 V x (flag[0] must be clear)
// ob: the element one-body arguments for the right-hand side (ndvr = 20),
// with its launch shape; null for the band kernel
void launch_cn_stage(const CnArgs& a, const OneBodyElArgs* ob, int shape, int lc, int mode, hipStream_t s);
// the convergence test of the term just enqueued (mode 1: kept set once set)
void launch_cn_check(const CnArgs& a, hipStream_t s);

struct PoissonArgs {
  int nr, n, nfe, ni, L2, npair;
  const int* pair_m;  // m_j - m_i of each pair
  const double *G, *Wl, *Wr, *Sd, *Sl, *Su;  // G: element inverses, transposed [L2][gstride][k][i]
  int gstride;                              // elements per L in G (nfe, or more in a shared table)
  const double *WlI, *WrI;                  // Wl, Wr times invr ([L2][nr])
  int pcr_steps;                            // cyclic reduction of the Schur system
  const double *pcrA, *pcrG, *pcrBinv;      // [L2][pcr_steps][nfe-1] x2, [L2][nfe-1]
  const double *invr, *rpl0, *rpl1, *fac2, *fac1;
  double2* Q;  // [npair][L2] multipole moments
};

struct MfArgs {
  int nr, nlat, nfun, nfcore, nfcore2, ncore, nact, nradfc, npair, nelact2;
  const int *pj, *pi, *ptype, *mval;
  const double2 *den1, *den2, *rden;
  // m-filtered coefficients written by the CI island: act2 [n^4] (den2
  // layout), actx weights -den1/2 [n^2], rden [n^2]
  const double2* mfc;
};

constexpr int kMaxGram = 2048;   // nL * nR entries of one Gram launch (bounds its partial-sum buffer)

struct GramArgs {
  size_t len;
  int nL, nR;
  int conj_left;        // 1: S = sum conj(L) R (inner products); 0: sum L R (zdotu)
  int with_conj_right;  // also accumulate sum opA(L) conj(R)
  const void* const* left;   // device array of nL device pointers (complex128, or complex64 when f32)
  const void* const* right;  // device array of nR device pointers
  int f32;                   // the vectors are complex64 (accumulation stays FP64)
};

struct DtorbArgs {
  int nbas, nr, nfun, nfcore;
  const int* mval;
  const double2 *h01, *gorb, *v2orb, *orb, *xmat;
  double2 tfac;
  // xlocal: each block builds its column of X from the l-space Gram products
  // (fc-dyn, core-core and active-active blocks; no core-active block, i.e.
  // ncore == nfcore or no active orbitals) instead of reading xmat
  int xlocal, ncore;
  // ladapt: 1 zero the derivative's l != lval[f] components of every orbital,
  // 2 of the cores only (clwfn::ladapt after the hprod)
  int ladapt;
  const int* lval;
  // split_type (hprod_htot_dtorb.F90:60-83): 0 the Q part carries v2 + g +
  // h1 + h0, 1 only v2 + g (h0 + h1 go to the clh1itr half steps)
  int isplit;
  double2 xfac;  // X = <j|H|i> * xfac
  const double2* lprod;
  const int *mapH, *mapG, *mapV, *mapF;
  // lnblk > 0: the l-space products as launch_gram's partial sums
  // (entry q at lpart[q * lnblk ...]), summed here instead of read from lprod
  const double2* lpart;
  int lnblk;
  double2* kout;          // stage derivative (may be null)
  double2* acc;           // psi += dtw * k (may be null)
  double dtw;
  const double* mask;     // applied to acc after the update (may be null)
  double2* next;          // next stage input (may be null)
  const double2* w0;      // step-start state (may alias acc: read before acc is written)
  double2* w0out;         // copy of acc as read, i.e. the step-start state (may be null)
  int nprev;
  const double2* kprev[kMaxStages];
  double cprev[kMaxStages];
  double cnew;
};

struct IslandArgs {
  int nfun, nfcore, ncore, nact, lcic, nelact2, icomp;
  const int* mval;
  double throcc1, throcc2, dtime;
  // CAS tables
  int nstr_a, nstr_b;
  const int *det_of, *exA, *sgA, *exB, *sgB;
  const double2* cic;
  double2* ci_work;       // (2 nact^2 + 2) * nstr_a * nstr_b; holds D_pq = E_pq C between the stages
  // dynamic LDS plan (island_lds_plan): CI work and string tables in LDS,
  // int2e in LDS at byte offset ci_lds_off
  int ci_lds, i2_lds;
  size_t ci_lds_off, lds_bytes;
  // stage A outputs
  double2 *den1, *den2, *rden, *rrden;
  // stage B inputs: Gram results and index maps ([nfun*nfun], -1 = absent)
  const double2* lprod;
  const int *mapH, *mapG, *mapV, *mapF;
  // lnblk > 0: the l-space products as launch_gram's partial sums
  // (entry q at lpart[q * lnblk ...]), summed here instead of read from lprod
  const double2* lpart;
  int lnblk;
  const double2* pprod;
  const int* map2e;        // [nact^4] (i,j,k,l) with i >= j, -1 otherwise
  // stage B outputs
  double2 *int1e, *int2e, *xmat, *dcic;
  double2* mfc;            // mean-field coefficients (see MfArgs)
  double* ene;             // ene_fcore (input), dcore, core, act, tot
  // Runge-Kutta stage update of the CI part, done after dcic (kout == null: none):
  // kout = dcic; acc += dtw * dcic; next = w0 + sum cprev[j] kprev[j] + cnew * dcic;
  // w0 may alias acc (read before acc is written); w0out (may be null) gets acc as read
  double2* kout;
  double2* acc;
  double dtw;
  double2* next;
  const double2* w0;
  double2* w0out;
  int nprev;
  const double2* kprev[kMaxStages];
  double cprev[kMaxStages];
  double cnew;
};

// The frozen-core phase of the velocity gauge (hprod_fcorbg,
// src/hprod/hprod_orbin.F90:68-105) as a job a transform launch can carry:
// orbg[f][th][r] = orbg0[f][th][r] exp(-i A r cos th) for r < nradfc, 0
// beyond; orbg has the transform's output type.
struct FcPhaseJob {
  const double2* orbg0;
  void* orbg;
  const double *xr, *cost, *lf;
  int nr, nradfc, nlat, nfc;
};

// path: 0 by grid size, 1 the vector form, 2 the matrix-core (MFMA) form
// tvar: launch shape of the register-blocked matrix-core form (0..5, same bits)
void launch_transform(int mode, const double2* in, double2* out, const double* mats, const TransformBatch* batch,
                      int nbatch, int nr, int K, int C, int Cp, double2* copy, hipStream_t s, int path = 0,
                      const FcPhaseJob* fc = nullptr, int tvar = 0);
inline int pad_cols(int C) { return (C + 31) / 32 * 32; }
// the same transform with complex64 on either side (matsf: the matrices in
// FP32, used when both sides are complex64)
template <class TI, class TO>
void launch_transform_t(int mode, const TI* in, TO* out, const double* mats, const float* matsf,
                        const TransformBatch* batch, int nbatch, int nr, int K, int C, int Cp, double2* copy,
                        hipStream_t s, int path = 0, const FcPhaseJob* fc = nullptr, int tvar = 0);
void launch_onebody(const double2* orb, double2* h01, const OneBodyArgs& a, hipStream_t s);
// shape 0/1/2: 3/6/12 elements per block; lc: l values per block (any
// shape gives the same bits; the autotuner picks one per problem)
bool launch_onebody_el(const double2* orb, double2* h01, const OneBodyElArgs& a, hipStream_t s, int shape = 2,
                       int lc = 2);
template <class T>  // rho and v complex128 or complex64; the solve is FP64
void launch_poisson(const T* rho, double2* xb, T* v, const PoissonArgs& a, hipStream_t s, int part = 0);  // part 1 bridge pass, 2 interior pass, 0 both
template <class T>  // theta-grid arrays complex128 or complex64 (arithmetic in the same)
void launch_mfprod(const T* orbg, const T* vg, T* gorbg, T* v2orbg, const MfArgs& a, hipStream_t s);
// out == nullptr: the partial sums only (partial[q * nblk + b], q = entry
// * 2 + c, nblk = gram_blocks(a, maxblk)), summed by the consumer
void launch_gram(const GramArgs& a, double2* partial, int maxblk, double2* out, hipStream_t s);
int gram_blocks(const GramArgs& a, int maxblk);
// Dipole expectation values: <z> = 2 sum_core <i|z|i> + sum_ab conj(den1(b,a)) <b|z|a>,
// and with nforms = 3 the same traces of p_z and a_z plus the field terms
struct DipoleArgs {
  int nr, L1, nfun, nfcore, ncore, nact, nradfc, mmax1;
  const int* mval;
  const double* xr;
  const double* alph;     // [2mmax1+1][L1]
  const double2* orb;     // orbitals as hprod_orbin builds them
  const double2* den1;    // [nact][nact]
  // nforms = 3 only
  int ndvr, igauge;
  double znuc, nel;       // nuclear charge, neltot(3)
  const double* band_p;   // [2n+1][nr] d/dr band
  const double* lf;       // device lfield[9]
};
// hprod_dipole's frozen-core corrections (src/hprod/hprod_dipole.F90:74-125).
// The generalised Fock rows of the frozen cores (hprod_gfock with the onFC
// branches of hprod_mfprod2_gfock): on the theta grid, rows r < nradfc,
// out[0][i] = g_i (dcj, dcx, fcx1 on the frozen core i), out[1][i] = v2_i
// (actj, actx on it).  The pair potentials are the hprod's (vg, with pj, pi,
// ptype as MfArgs), the frozen-core pairs vfc[(i nfc + j) ng + x] = v(j, i)
// (the potential of conj(phi_j) phi_i).
struct GfockFcArgs {
  int nr, nlat, nradfc, nfun, nfcore, nfcore2, ncore, nact, npair;
  const int *pj, *pi, *ptype, *mval;
  const double2* den1;
  const double2* vfc;
};
template <class T>
void launch_gfock_fc(const T* orbg, const T* vg, T* out, const GfockFcArgs& a, hipStream_t s);
// The correction sums over the rows r < nradfc and all l, with
// f_x = h01_x + g_x + v2_x (+ E_z z phi_x for the dynamical orbitals in the
// velocity gauge; the frozen cores' h01 carries it already), Z = z phi,
// P = p_z phi at zfac = 1: q = 0: sum_fc <Z_i|f_i>, 1: sum_fc <P_i|f_i>,
// 2 + i nf + x: <phi_i|f_x> (i frozen, m_i = m_x), 2 + nfc nf + x nfc + i:
// <Z_x|phi_i>, 2 + 2 nfc nf + x nfc + i: <P_x|phi_i>.  The final kernel
// forms vel(2), acc(2) from them and writes {dip, vel, acc} x {(1), (2),
// (3)} with the (1) terms read from d3.
struct DipoleFcArgs {
  int nr, L1, nfun, nfcore, ncore, nact, nradfc, mmax1, ndvr, igauge;
  const int* mval;
  const double *xr, *alph, *band_p, *lf;
  const double2 *orb, *h01, *gdyn, *vdyn, *gfc, *vfc;  // gdyn, vdyn: [nfun - nfcore][L1][nr]; gfc, vfc: [nfcore][L1][nr]
  const double2* den1;
};
int dipole_fc_quantities(const DipoleFcArgs& a);
int dipole_fc_blocks(const DipoleFcArgs& a);  // per quantity
void launch_dipole_fc(const DipoleFcArgs& a, double2* partial, const double* d3, double* out9, hipStream_t s);
// out9 = {d3[f], 0, d3[f]} per form f (no frozen core: no corrections)
void launch_dipole_expand(const double* d3, double* out9, hipStream_t s);
// partial: nforms * dipole_blocks(a) entries
int dipole_blocks(const DipoleArgs& a);
void launch_dipole(const DipoleArgs& a, int nforms, double2* partial, double* out, hipStream_t s);
// ovlp[j + nf*i] = <orb_j|orb_i> over the radial rows r < cut (m-matched), nf*nf entries
void launch_window_ovlp(const double2* orb, const int* mval, int nf, int nr, int L1, int cut, double2* ovlp,
                        hipStream_t s);
// clwfn::orth (schmidt): CI normalised (lcic > 0), dynamical orbitals by modified Gram-Schmidt
void launch_orth(double2* orb, double2* cic, const int* mval, int nr, int L1, int nfcore, int nfun, int lcic,
                 hipStream_t s);
void launch_to_f32(float2* y, const double2* x, size_t n, hipStream_t s);
void launch_zero(double2* y, size_t n, hipStream_t s);
// z = a x + b y over n complex values; a null x or y drops its term
void launch_zaxpby(size_t n, double2 a, const double2* x, double2 b, const double2* y, double2* z, hipStream_t s);
// the six field arrays of cletdrb step st->n (t = t0 + n dt), then n += 1:
// lfield(t), lfield(t + dt/2), lfield(t + dt), d lfield/dt (t),
// lfield(t + dt/2) - lfield(t), lfield(t + dt) - lfield(t)
void launch_etd_fields(const Pulse& p, PropClock* st, double* lf, hipStream_t s);
// clexpab's field arrays: lfield(t), then lfield(tp[k]) - lfield(t), k = 1..3
void launch_xab_fields(const Pulse& p, double t, const double* tp, double* lf, hipStream_t s);
void launch_field_stages(const Pulse& p, PropClock* st, const StageNodes& nodes, double* lf, hipStream_t s);
void launch_clock_set(PropClock* st, double t0, double dt, hipStream_t s);
void launch_dtorb(const DtorbArgs& a, hipStream_t s);
void launch_mask(double2* wfn, const double* mask, int nr, int L1, int nfcore, int nfun, hipStream_t s);
constexpr size_t kIslandLds = 32768;  // dynamic LDS budget of the island kernels
void island_lds_plan(IslandArgs& a);
void launch_island_a(const IslandArgs& a, hipStream_t s);
void launch_island_b(const IslandArgs& a, hipStream_t s);

// ---------------------------------------------------------------------------
// Batches of independent problems (scans over pulses, grids and lmax).  While
// a recorder is active on the calling thread, the launch functions above
// append what they would launch (kind, template variant, grid, block,
// dynamic LDS and the kernel's arguments) to it instead of launching.  A
// batch records one step of every member, then issues one launch per
// recorded site over all members: block b of such a launch belongs to the
// member m with start[m] <= b < start[m + 1], takes its coordinates within
// that member's own grid from dims[3m .. 3m + 2], and runs the kernel body
// on the member's arguments recs[m], as the member's own launch would.
// ---------------------------------------------------------------------------
struct Blk {
  int x, y, z, nx, ny, nz;
};
__device__ __forceinline__ Blk hw_blk() {
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}
struct BatchMap {
  const int* start;  // [nmem + 1] first block of each member
  const int* dims;   // [nmem][3] the member's grid
  int nmem;
  // 0: hardware block b walks blocks b, b + gridDim.x, ... of the whole
  // launch; k > 0: hardware block b serves member b / k alone and walks its
  // blocks b % k, b % k + k, ... (gridDim.x = nmem * k)
  int per_member;
};
// Walks the blocks of a batched launch: block b of member m (start[m] <= b <
// start[m + 1]) is run as f(m, its coordinates in m's grid) by hardware
// block b mod gridDim.x, in order, with a barrier between two runs.  A launch
// with fewer hardware blocks than blocks walks them grid-stride, and the
// blocks of a member for which skip(m) holds are passed over in O(1) (the
// half-step series of a converged member).
template <class Skip, class F>
__device__ __forceinline__ void batch_walk(const BatchMap& map, Skip skip, F f) {
  if (map.per_member > 0) {
    const int k = map.per_member;
    const int m = (int)blockIdx.x / k;
    if (m >= map.nmem || skip(m)) return;
    const int beg = map.start[m], n = map.start[m + 1] - beg;
    const int nx = map.dims[3 * m], ny = map.dims[3 * m + 1], nz = map.dims[3 * m + 2];
    bool first = true;
    for (int q = (int)blockIdx.x - m * k; q < n; q += k) {
      if (!first) __syncthreads();
      first = false;
      f(m, Blk{q % nx, (q / nx) % ny, q / (nx * ny), nx, ny, nz});
    }
    return;
  }
  const int total = map.start[map.nmem];
  const int G = (int)gridDim.x;
  int m = 0, m_beg = 0, m_end = -1;
  bool first = true;
  for (int vb = (int)blockIdx.x; vb < total; vb += G) {
    if (vb >= m_end) {
      int lo = 0, hi = map.nmem;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (map.start[mid] <= vb) lo = mid;
        else hi = mid;
      }
      m = lo;
      m_beg = map.start[m];
      m_end = map.start[m + 1];
      if (skip(m)) {
        // the first block of this hardware block past member m, less one stride
        vb += ((m_end - vb + G - 1) / G - 1) * G;
        continue;
      }
    }
    if (!first) __syncthreads();
    first = false;
    const int q = vb - m_beg;
    const int nx = map.dims[3 * m], ny = map.dims[3 * m + 1], nz = map.dims[3 * m + 2];
    f(m, Blk{q % nx, (q / nx) % ny, q / (nx * ny), nx, ny, nz});
  }
}
struct NoSkip {
  __device__ bool operator()(int) const { return false; }
};
// Member m's argument record, copied into the block's LDS buffer `raw` when
// the member changes (have: the member held).  The kernel body then reads
// its arguments from LDS, which its stores to global memory cannot alias, so
// they stay in registers or cheap LDS reads instead of being reloaded from
// global memory after every store.
template <class R>
__device__ __forceinline__ const R& stage_rec(const R* __restrict__ recs, int m, unsigned char* raw, int& have) {
  static_assert(sizeof(R) % 4 == 0, "argument records are whole words");
  if (m != have) {
    const int* src = reinterpret_cast<const int*>(recs + m);
    int* dst = reinterpret_cast<int*>(raw);
    for (int i = (int)threadIdx.x; i < (int)(sizeof(R) / 4); i += (int)blockDim.x) dst[i] = src[i];
    __syncthreads();
    have = m;
  }
  return *reinterpret_cast<const R*>(raw);
}

enum RecKind {
  kRecTransformW = 1,   // var: CT, U, type code
  kRecTransformM,       // var: MODE, type code (256 threads)
  kRecOneBody,          // var: GE, CN
  kRecPoissonBridge,    // var: type code (LDS form, ndvr = 20)
  kRecPoissonInterior,  // var: type code
  kRecMfprodT,          // var: NF, type code (dynamical cores only)
  kRecMfprod,           // var: type code
  kRecGram,             // var: TA, TB, CR, f32
  kRecDtorb,
  kRecMask,
  kRecField,
  kRecIslandB,
  kRecCnBridge,
  kRecCnInterior,
  kRecCnCheck,
};
// type codes of (input, output): 0 (c128, c128), 1 (c128, c64), 2 (c64, c64), 3 (c64, c128)
struct TransformRec {
  const void* in;
  void* out;
  const double* mats;
  const float* matsf;
  const TransformBatch* batch;
  int nr, K, C, Cp;
  double2* copy;
  int nx, ny, ntile;
  FcPhaseJob fc;
};
// In a batch, the terms of the half-step series share one word: the check
// of term k stores k when some member has not converged, and a launch of
// term k > 1 (right-hand side, bridge, interior, check) runs only when the
// word holds k - 1, so that a term after the last member has converged
// costs one load per block.  The first term's bridge pass clears it.
struct SeriesGate {
  int* word;  // null outside a batch
  int term;   // the term this launch forms (1: the first)
};
// true when the launch is a later term of a series that has ended
__device__ __forceinline__ bool gate_closed(const SeriesGate& g) {
  return g.word && g.term > 1 && *g.word != g.term - 1;
}
struct OneBodyRec {
  const double2* orb;
  double2* h01;
  OneBodyElArgs a;
  int lc;
  SeriesGate gate;  // cn = 2 launches
};
struct PoissonRec {
  const void* rho;
  void* v;
  double2* xb;
  PoissonArgs a;
  int epb, nchunk;
};
struct MfRec {
  const void *orbg, *vg;
  void *gorbg, *v2orbg;
  MfArgs a;
};
struct GramRec {
  GramArgs a;
  double2* partial;
  int nqblk;
};
struct MaskRec {
  double2* wfn;
  const double* mask;
  int nr, L1, nfcore, nfun;
};
struct FieldRec {
  Pulse p;
  PropClock* st;
  StageNodes nodes;
  double* lf;
};
struct CnRec {
  CnArgs a;
  int mode, epb, nchunk;
  SeriesGate gate;
};

struct LaunchRec {
  int kind = 0;
  int var[4] = {0, 0, 0, 0};
  int grid[3] = {1, 1, 1};
  int block = 0;
  size_t lds = 0;
  std::vector<unsigned char> args;
};
struct Recorder {
  std::vector<LaunchRec> recs;
  std::string error;  // the first launch that has no batched form
};
// the calling thread's active recorder (null: launch normally)
Recorder*& active_recorder();
template <class R>
void record_launch(int kind, const int (&var)[4], dim3 grid, int block, size_t lds, const R& rec) {
  LaunchRec L;
  L.kind = kind;
  for (int i = 0; i < 4; ++i) L.var[i] = var[i];
  L.grid[0] = (int)grid.x;
  L.grid[1] = (int)grid.y;
  L.grid[2] = (int)grid.z;
  L.block = block;
  L.lds = lds;
  const unsigned char* b = reinterpret_cast<const unsigned char*>(&rec);
  L.args.assign(b, b + sizeof(R));
  active_recorder()->recs.push_back(std::move(L));
}
inline void record_unsupported(const char* what) {
  Recorder* r = active_recorder();
  if (r->error.empty()) r->error = what;
}
// One batched launch of site `p` over nmem members: recs holds nmem
// argument records of the site's kind, map the block ranges; `grid`
// hardware blocks (the blocks of all members, or fewer for a site that
// batch_skips() says skips converged members).
// Returns false when the kind or variant has no batched kernel in that
// file; dry: only report whether it has one.
bool launch_batch_kernels(const LaunchRec& p, const void* recs, const BatchMap& map, int grid, size_t lds,
                          hipStream_t s, bool dry);
bool launch_batch_h1itr(const LaunchRec& p, const void* recs, const BatchMap& map, int grid, size_t lds,
                        hipStream_t s, bool dry);
bool launch_batch_island(const LaunchRec& p, const void* recs, const BatchMap& map, int grid, size_t lds,
                         hipStream_t s, bool dry);

// The Poisson interior pass of a batch whose members share one element
// geometry (equal element widths, orders and charges-free operator): the
// element inverses G_e and the row tables at (L, point) are then the same
// for every member and are read from one copy of the largest extent.
// Items are (member, pair) pairs in groups of kPoissonSlots (member -1: an
// empty slot); a job is (group, L, element chunk) and runs as one block.
constexpr int kPoissonSlots = 8;
constexpr int kPoissonInteriorThreads = 256;  // threads of the interior passes
struct PoissonShared {
  const double *G, *Wl, *Wr, *invr, *rpl1, *fac2;
  int gstride;  // elements per L in G
  int rstride;  // points per L in the row tables
  int epb;      // elements per block (every member's)
};
// type: the site's type code (0: complex128, 2: complex64); false for others
bool launch_poisson_interior_grouped(int type, const void* recs, const int2* items, const int* jobs, int njobs,
                                     const PoissonShared& sh, hipStream_t s);
// The interior pass of the clh1itr half step in the same form: items are
// (member, dynamical orbital) pairs in groups of kCnGroup over the shared
// element inverses; each member must have one dynamical orbital (its block
// norm is then the item's, written where the member's own pass writes it).
struct CnShared {
  const double2* G;
  int gstride;  // elements per l in G
  int epb;      // elements per block (every member's)
  SeriesGate gate;
};
void launch_cn_interior_grouped(const void* recs, const int2* items, const int* jobs, int njobs, const CnShared& sh,
                                hipStream_t s);
// sites whose blocks return at once for a converged member (the later terms
// of the half-step series): launched member-major (BatchMap::per_member), so
// that a converged member costs one flag read per hardware block
inline bool batch_skips(const LaunchRec& p) {
  return ((p.kind == kRecCnBridge || p.kind == kRecCnInterior || p.kind == kRecCnCheck) && p.var[0] == 1) ||
         (p.kind == kRecOneBody && p.var[1] == 2);
}
// the argument record size of a kind
size_t rec_size(int kind);

}  // namespace td1c
