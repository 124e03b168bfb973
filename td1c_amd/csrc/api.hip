// C-ABI of the gfx950 hprod (include/td1c_hprod.h): context creation,
// frozen-core set-up, the hprod launch sequence, the Runge-Kutta step, the
// mask and the laser field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/td1c_hprod.h"
#include "field.hpp"
#include "kernels.hpp"
#include "setup.hpp"

namespace td1c {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(-3, std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(&p, count * sizeof(T));
  }
  hipError_t upload(const std::vector<T>& v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

// The side streams of a device, created once per process and shared by all
// contexts on it.  HIP maps streams onto a few hardware queues in creation
// order; one fixed pair keeps every context's branches on the same queues
// (per-context streams made the mapping, and the run time, depend on how
// many contexts had been created before).
hipError_t side_streams(int device, hipStream_t out[2]) {
  static std::mutex mu;
  static std::map<int, std::pair<hipStream_t, hipStream_t>> pool;
  std::lock_guard<std::mutex> lock(mu);
  auto it = pool.find(device);
  if (it == pool.end()) {
    hipStream_t a = nullptr, b = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    it = pool.emplace(device, std::make_pair(a, b)).first;
  }
  out[0] = it->second.first;
  out[1] = it->second.second;
  return hipSuccess;
}

}  // namespace td1c

using namespace td1c;

struct td1c_ctx {
  Setup S;
  int npair = 0, nd = 0, nbas = 0, ngrid = 0;
  int device = 0;
  // tables
  DevBuf<double> xr, band_k, band_k0, band_p, alph, cost, mask;
  DevBuf<double2> tdiag, tzor;
  DevBuf<double> eK0, eK, eP;
  DevBuf<double> m_l2g, m_g2l, m_g2L, m_L2g;
  DevBuf<double> pG, pWl, pWr, pWlI, pWrI, pcrA, pcrG, pcrBinv, pSd, pSl, pSu, invr, rpl0, rpl1, fac2, fac1;
  DevBuf<int> mval, lval, pair_m, pj, pi, ptype, emap;
  // batches
  DevBuf<TransformBatch> b_l2g_dyn, b_g2l_fc, b_rho, b_L2g, b_g2l_mf;
  int n_l2g_dyn = 0, n_g2l_fc = 0, n_g2l_mf = 0;
  // work space
  DevBuf<double2> lsp;       // orb | h01 | gorb | v2orb
  double2 *orb = nullptr, *h01 = nullptr, *gorb = nullptr, *v2orb = nullptr;
  DevBuf<double2> gsp;       // orbg | gorbg | v2orbg
  double2 *orbg = nullptr, *gorbg = nullptr, *v2orbg = nullptr;
  // mixed precision (precision = 1): the theta-grid and pair arrays in complex64
  bool mixed = false;
  DevBuf<float2> gsp_f, rhoL_f, vL_f, vg_f;
  float2 *orbg_f = nullptr, *gorbg_f = nullptr, *v2orbg_f = nullptr;
  DevBuf<float> m_g2L_f, m_L2g_f;
  DevBuf<double2> orbg0;     // frozen-core theta-grid orbitals at A = 0
  DevBuf<double2> rhoL, vL, vg;
  DevBuf<double2> p_xb, p_Q;  // Poisson bridge solutions [pair][L][nfe-1], multipole moments [pair][L]
  // Gram
  DevBuf<const void*> g_left_l, g_right_l, g_left_p, g_right_p;
  int gl_nL = 0, gl_nR = 0, gp_nL = 0, gp_nR = 0, gram_blk = 0;
  DevBuf<double2> gram_partial, gram_partial_p, lprod, pprod;
  DevBuf<int> mapH, mapG, mapV, mapF, map2e;
  // island
  DevBuf<double2> den1, den2, rden, rrden, int1e, int2e, xmat, dcic, ci_work, mfc, dip_part, ovlp;
  DevBuf<double> ene;
  DevBuf<int> det_of, exA, sgA, exB, sgB;
  double ene_fcore = 0.0;
  // RK work: w0, t, k[stages] over the full wave function
  DevBuf<double2> rk_w0, rk_t, rk_k;
  DevBuf<double> lf_dev;  // field values of the stages, [kMaxStages][9]
  DevBuf<PropClock> clock;
  // split_type = 1 (clh1itr): Crank-Nicolson factors for half steps of
  // cn_dt / 2, the series terms and the convergence flag
  double cn_dt = 0.0;
  int cn_maxcyc = 20;
  int* cn_flag_host = nullptr;  // pinned: the series' flag read back by stream launches
  double cn_thresh = 1.0e-10;
  int cn_K = 0, cn_npart1 = 0;
  DevBuf<double2> cnG, cnWl, cnWr, cnPa, cnPg, cnBinv, cn_x, cn_y;
  DevBuf<double> cn_part, lf_cn;
  DevBuf<int> cn_flag;
  std::vector<double> v2jfc_h;  // frozen-core Hartree potential (complex, host), empty without frozen core
  std::vector<double> cnG_h;    // host copy of cnG (complex, [L1][nfe][k][i])
  // frozen-core dipole corrections (td1c_dipole_all): the frozen-core pair
  // potentials v(j, i) on the theta grid, the frozen cores' generalised Fock
  // rows (theta grid, then l-space: g, v2), their theta -> l batch, partials
  DevBuf<double2> vfc_tab, gfc_g, gfc_l, dipfc_part;
  DevBuf<float2> gfc_gf;
  DevBuf<TransformBatch> b_g2l_gfc;
  DevBuf<double> dip3_tmp;
  // device copies of hprod_htot_'s host arguments (wfn | hwfn | cic | hcic),
  // allocated by its first call and kept with the context
  DevBuf<double2> legacy_stage;
  bool wfn0_set = false;
  // two side streams (default priority) for the independent branches of an
  // hprod, joined back by events; the events carry no timing.  The streams
  // belong to the process (side_streams) and are shared by every context on
  // the device.
  hipStream_t side[2] = {nullptr, nullptr};
  // fork/join events: [0] pair potentials ready, [1] RDM stage done,
  // [2] orbitals on the theta grid, [3] one-body part done, [4] pair Gram
  // products done, [5] l-space Gram products done, [6] island stage B done
  // (when it runs on side stream 0)
  hipEvent_t fj[7] = {};
  // X built by the orbital-derivative kernel (no core-active block), so the
  // island stage B runs on side stream 0, off the main stream
  bool x_local = false;
  int tpath = 0;  // transform form: 0 by grid size, 1 vector, 2 matrix cores
  // launch shape of the element one-body kernel (autotuned per problem):
  // elements per block (0: 3, 1: 6, 2: 12) and l values per block
  int ob_shape = 2, ob_lc = 2;
  // launch shape of the register-blocked matrix-core transforms (0..5;
  // TD1C_TVAR in the environment overrides it): 1 = 2 column tiles per
  // wavefront, 4 contraction steps per load batch, measured best on the Ar
  // workload among (4,4) (2,4) (4,8) (8,4) (8,2) (2,8)
  int tvar = 1;
  std::string tune_log;  // what the autotuner measured (JSON)
  // device-driven propagation: one captured RK step, replayed per step
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool use_graph = false;  // td1c_set_graph
  bool capturing = false;
  int g_order = 0, g_mask = 0;
  double g_dt = 0.0;
  const double* g_wfn = nullptr;
  Pulse g_pulse{};
  // optional per-kernel timing with HIP events on the launch stream
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  // cletdrb (td1c_etdrb_gen): the seven rational functions of h1 of
  // prop4_orb, the element Schur tables of their distinct denominator
  // factors cd0 + cd1 T (as 1 + (cd1/cd0) T), the step's states and fields
  struct EtdTab {
    DevBuf<double2> G, Wl, Wr, pa, pg, binv;
    int K = 0;
    double fre = 0.0, fim = 0.0;
  };
  struct EtdRat {
    double dt = 0.0;
    int dimn = 0, dimd = 0;
    double2 cn[3], cd0[3], cd1[3];
    double2 numer_limit, denom_limit, coeff_limit;
    double2 rk;  // cletdrk4's CI factor of this function
    int tab[3];
  };
  std::vector<std::unique_ptr<EtdTab>> etd_tabs;
  EtdRat etd[7];
  int etd_scheme = 0;  // 1 cletdrb, 2 cletdrk4, 3 cllawrk, 4 clexpab
  // clexpab: the inverse Taylor matrix (column-major), steps since gen, the
  // times of the history states
  double xab_ainv[16];
  int xab_nstep = 0;
  double xab_times[4] = {0.0, 0.0, 0.0, 0.0};
  bool etd_ready = false;
  int etd_rosen = 1, etd_maxcyc = 20;
  double etd_dt = 0.0, etd_thresh = 1.0e-10;
  DevBuf<double2> etd_buf;  // 13 states, then orb (orbital block) and three dynamical blocks
  // hprod_chkconv: the orbitals then the generalised Fock vectors of the
  // dynamical ones, the CI derivative; their overlaps; the combined m values
  DevBuf<double2> chk_buf, chk_ov;
  DevBuf<int> chk_mval;
  DevBuf<double> etd_lf;    // [6][9] field arrays of the step
  DevBuf<int> etd_zero;     // an always-clear stop flag

  struct Rec { const char* name; int e0, e1; };
  std::vector<Rec> recs;
  int ev_next = 0;
  std::string timing_only;  // empty: every kernel
  std::vector<std::pair<std::string, std::pair<double, int>>> timing_agg;

  ~td1c_ctx() {
    DevBuf<double>* d[] = {&eK0, &eK, &eP, &xr, &band_k, &band_k0, &band_p, &alph, &cost, &mask, &m_l2g, &m_g2l, &m_g2L, &m_L2g,
                           &pG, &pWl, &pWr, &pWlI, &pWrI, &pcrA, &pcrG, &pcrBinv, &pSd, &pSl, &pSu, &invr, &rpl0, &rpl1, &fac2, &fac1, &ene, &lf_dev,
                       &cn_part, &lf_cn};
    for (auto* b : d) b->release();
    DevBuf<double2>* z[] = {&tdiag, &tzor, &lsp, &gsp, &orbg0, &rhoL, &vL, &vg, &p_xb, &p_Q, &legacy_stage,
                            &cnG, &cnWl, &cnWr, &cnPa, &cnPg, &cnBinv, &cn_x, &cn_y, &vfc_tab, &gfc_g, &gfc_l,
                            &dipfc_part,
                            &gram_partial, &gram_partial_p, &lprod, &pprod, &den1, &den2, &rden, &rrden, &int1e, &int2e, &xmat,
                            &dcic, &ci_work, &mfc, &dip_part, &ovlp, &rk_w0, &rk_t, &rk_k};
    for (auto* b : z) b->release();
    DevBuf<int>* i[] = {&mval, &lval, &pair_m, &pj, &pi, &ptype, &emap, &mapH, &mapG, &mapV, &mapF, &map2e, &cn_flag,
                        &det_of, &exA, &sgA, &exB, &sgB};
    for (auto* b : i) b->release();
    DevBuf<TransformBatch>* t[] = {&b_l2g_dyn, &b_g2l_fc, &b_rho, &b_L2g, &b_g2l_mf};
    for (auto* b : t) b->release();
    DevBuf<const void*>* g[] = {&g_left_l, &g_right_l, &g_left_p, &g_right_p};
    for (auto* b : g) b->release();
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    DevBuf<float2>* ff[] = {&gsp_f, &rhoL_f, &vL_f, &vg_f, &gfc_gf};
    b_g2l_gfc.release();
    dip3_tmp.release();
    for (auto* b : ff) b->release();
    DevBuf<float>* fm[] = {&m_g2L_f, &m_L2g_f};
    for (auto* b : fm) b->release();
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (cn_flag_host) (void)hipHostFree(cn_flag_host);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto& e : fj)
      if (e) (void)hipEventDestroy(e);
    if (clock.p) (void)hipFree(clock.p);
    for (auto& t : etd_tabs) {
      DevBuf<double2>* e[] = {&t->G, &t->Wl, &t->Wr, &t->pa, &t->pg, &t->binv};
      for (auto* b : e) b->release();
    }
    etd_buf.release();
    chk_buf.release();
    chk_ov.release();
    chk_mval.release();
    etd_lf.release();
    etd_zero.release();
  }
};

// A batch of independent problems propagated together (td1c_batch_*): the
// members' own contexts hold every table and work array; the batch holds one
// recorded step (a list of launch sites, each with one argument record per
// member) in a device table, and issues one launch per site for all members.
struct td1c_batch {
  std::vector<td1c_ctx*> mem;
  int device = 0;
  bool use_graph = true;
  int ob_shape = 2, ob_lc = 2;  // element one-body launch shape of the batch
  int tvar = 1;                 // register-blocked transform shape of the batch
  struct Site {
    LaunchRec proto;  // kind, variant and block of every member's launch
    int total = 0;    // blocks over all members
    int max_blocks = 0;  // blocks of the largest member
    size_t lds = 0;
    size_t off_recs = 0, off_start = 0, off_dims = 0;
    // interior passes over groups of items reading the shared tables:
    // 1 the Poisson solve (launch_poisson_interior_grouped), 2 the clh1itr
    // half step (launch_cn_interior_grouped)
    int grouped = 0, njobs = 0;
    size_t off_items = 0, off_jobs = 0;
    PoissonShared psh{};
    CnShared csh{};
  };
  std::vector<Site> sites;
  DevBuf<unsigned char> table;
  size_t table_bytes = 0;
  // element tables shared by the members whose own tables equal them
  // (equal element boundaries, charges and dt give equal element blocks):
  // Crank-Nicolson and Poisson interior inverses, element one-body matrices
  DevBuf<double2> u_cnG;
  DevBuf<double> u_pG, u_eK0, u_eK, u_eP;
  // the Poisson row tables (Wl, Wr, invr, rpl1, fac2) at [L][point] with
  // rstride points per L, when every member's equal the shared copy
  DevBuf<double> u_prow[5];
  int rstride = 0;
  int shared[4] = {0, 0, 0, 0};  // members reading each shared table (3: the Poisson row tables)
  int pgroups = 0, cgroups = 0;   // item groups of the grouped Poisson and half-step interior passes
  DevBuf<int> cn_word;            // the half-step series' word (SeriesGate)
  std::vector<unsigned char> key;  // what the recorded step depends on
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::string tune_log;
  ~td1c_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    table.release();
    u_cnG.release();
    u_pG.release();
    u_eK0.release();
    u_eK.release();
    u_eP.release();
    for (auto& t : u_prow) t.release();
    cn_word.release();
  }
};

namespace {

td1c_ctx* g_bound = nullptr;

// Folds the completed event pairs into the per-kernel totals and frees the pool.
hipError_t harvest_timing(td1c_ctx* c) {
  if (c->recs.empty()) return hipSuccess;
  hipError_t e = hipEventSynchronize(c->ev_pool[c->recs.back().e1]);
  if (e != hipSuccess) return e;
  for (auto& r : c->recs) {
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, c->ev_pool[r.e0], c->ev_pool[r.e1]);
    if (e != hipSuccess) return e;
    auto it = std::find_if(c->timing_agg.begin(), c->timing_agg.end(), [&](auto& x) { return x.first == r.name; });
    if (it == c->timing_agg.end()) c->timing_agg.push_back({r.name, {ms, 1}});
    else { it->second.first += ms; it->second.second += 1; }
  }
  c->recs.clear();
  c->ev_next = 0;
  return hipSuccess;
}

// Records an event pair around one launch when timing is on (for every
// kernel, or only for the one named by td1c_set_timing_kernel).  A full event
// pool is folded into the totals first, which waits for the recorded work.
struct Timed {
  td1c_ctx* c;
  hipStream_t s;
  const char* name;
  int e0 = -1;
  Timed(td1c_ctx* c_, hipStream_t s_, const char* n) : c(c_), s(s_), name(n) {
    if (!c->timing || c->capturing) return;
    if (!c->timing_only.empty() && c->timing_only != name) return;
    if (c->ev_next + 2 > (int)c->ev_pool.size() && harvest_timing(c) != hipSuccess) return;
    e0 = c->ev_next;
    c->ev_next += 2;
    (void)hipEventRecord(c->ev_pool[e0], s);
  }
  ~Timed() {
    if (e0 < 0) return;
    (void)hipEventRecord(c->ev_pool[e0 + 1], s);
    c->recs.push_back({name, e0, e0 + 1});
  }
};
#define TIMED(name) Timed timed_##__LINE__(c, s, name)
#define TIMED_ON(stream, name) Timed timed_##__LINE__(c, stream, name)

template <class T>
std::vector<T> flat_or_one(const std::vector<T>& v) {
  return v.empty() ? std::vector<T>(1, T()) : v;
}

// [nm][K][C] -> [nm][K][Cp] with the columns padded by zeros to a multiple of 32
std::vector<double> padded_cols(const std::vector<double>& m, int K, int C) {
  const int Cp = pad_cols(C);
  const size_t nm = m.size() / ((size_t)K * C);
  std::vector<double> out(nm * K * Cp, 0.0);
  for (size_t i = 0; i < nm; ++i)
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < C; ++j) out[(i * K + k) * Cp + j] = m[(i * K + k) * C + j];
  return out;
}

OneBodyElArgs onebody_el_args(td1c_ctx* c, const double* d_lf) {
  const Setup& S = c->S;
  OneBodyElArgs oe{};
  oe.nr = S.nr; oe.nfe = S.nfe; oe.L1 = S.L1; oe.ndvr = S.n; oe.nfcore = S.nfcore; oe.nradfc = S.nradfc;
  oe.nfun = S.nfun; oe.mmax1 = S.mmax1; oe.igauge = S.igauge; oe.fc_full = S.iprojfc == 2; oe.mval = c->mval.p;
  oe.eK0 = c->eK0.p; oe.eK = c->eK.p; oe.eP = c->eP.p; oe.emap = c->emap.p; oe.xr = c->xr.p; oe.alph = c->alph.p; oe.tzor = c->tzor.p;
  oe.lf = d_lf;
  return oe;
}

int build_device(td1c_ctx* c) {
  Setup& S = c->S;
  const int nr = S.nr, nf = S.nfun, nfc = S.nfcore, L1 = S.L1, L2 = S.L2, nl = S.nlat;
  c->nd = S.ndyn;
  c->nbas = nr * L1;
  c->ngrid = nr * nl;
  c->npair = (int)S.pairs.size();
  HIP_TRY(hipGetDevice(&c->device));
  HIP_TRY(c->xr.upload(S.xr));
  HIP_TRY(c->band_k.upload(S.bandT_k));
  HIP_TRY(c->band_k0.upload(S.bandT_k0));
  HIP_TRY(c->band_p.upload(S.bandT_p));
  HIP_TRY(c->alph.upload(S.alph));
  HIP_TRY(c->cost.upload(S.cost));
  {
    std::vector<double> m(nr);
    for (int i = 0; i < nr; ++i) m[i] = S.mask[i + 1];
    HIP_TRY(c->mask.upload(m));
  }
  {
    std::vector<double2> td((size_t)L1 * nr), tz((size_t)L1 * nr);
    for (size_t q = 0; q < td.size(); ++q) {
      td[q] = make_double2(S.tdiag[2 * q], S.tdiag[2 * q + 1]);
      tz[q] = make_double2(S.tzor[2 * q], S.tzor[2 * q + 1]);
    }
    HIP_TRY(c->tdiag.upload(td));
    HIP_TRY(c->tzor.upload(tz));
  }
  HIP_TRY(c->eK0.upload(S.eK0));
  HIP_TRY(c->eK.upload(S.eK));
  HIP_TRY(c->eP.upload(S.eP));
  HIP_TRY(c->emap.upload(S.emap));
  HIP_TRY(c->m_l2g.upload(padded_cols(S.mat_l2g, S.L1, S.nlat)));
  HIP_TRY(c->m_g2l.upload(padded_cols(S.mat_g2l, S.nlat, S.L1)));
  HIP_TRY(c->m_g2L.upload(padded_cols(S.mat_g2L, S.nlat, S.L2)));
  HIP_TRY(c->m_L2g.upload(padded_cols(S.mat_L2g, S.L2, S.nlat)));
  {
    // element inverses G_e stored transposed on the device, [L][e][k][i]:
    // lanes owning consecutive rows i read consecutive addresses
    const int ni = S.ni;
    std::vector<double> gt(S.pG.size());
    for (size_t blk = 0; blk < S.pG.size() / ((size_t)ni * ni); ++blk)
      for (int i = 0; i < ni; ++i)
        for (int k = 0; k < ni; ++k) gt[blk * ni * ni + (size_t)k * ni + i] = S.pG[blk * ni * ni + (size_t)i * ni + k];
    HIP_TRY(c->pG.upload(gt));
  }
  HIP_TRY(c->pWl.upload(S.pWl));
  HIP_TRY(c->pWr.upload(S.pWr));
  {
    // W with 1/r folded in: the bridge pass takes W . (rho invr) from raw rho
    std::vector<double> wl(S.pWl.size()), wr(S.pWr.size());
    for (size_t k = 0; k < wl.size(); ++k) {
      wl[k] = S.pWl[k] * S.d2invr[k];
      wr[k] = S.pWr[k] * S.d2invr[k];
    }
    HIP_TRY(c->pWlI.upload(wl));
    HIP_TRY(c->pWrI.upload(wr));
  }
  HIP_TRY(c->pcrA.upload(S.pcrA));
  HIP_TRY(c->pcrG.upload(S.pcrG));
  HIP_TRY(c->pcrBinv.upload(S.pcrBinv));
  HIP_TRY(c->pSd.upload(S.pSd));
  HIP_TRY(c->pSl.upload(S.pSl));
  HIP_TRY(c->pSu.upload(S.pSu));
  HIP_TRY(c->invr.upload(S.d2invr));
  HIP_TRY(c->rpl0.upload(S.d2rpl0));
  HIP_TRY(c->rpl1.upload(S.d2rpl1));
  HIP_TRY(c->fac2.upload(S.d2fac2));
  HIP_TRY(c->fac1.upload(S.d2fac1));
  HIP_TRY(c->mval.upload(S.mval));
  HIP_TRY(c->lval.upload(S.lval));
  {
    std::vector<int> pm, pj, pi, pt;
    for (auto& p : S.pairs) {
      pm.push_back(S.mval[p.i] - S.mval[p.j]);
      pj.push_back(p.j);
      pi.push_back(p.i);
      pt.push_back(p.type);
    }
    HIP_TRY(c->pair_m.upload(flat_or_one(pm)));
    HIP_TRY(c->pj.upload(flat_or_one(pj)));
    HIP_TRY(c->pi.upload(flat_or_one(pi)));
    HIP_TRY(c->ptype.upload(flat_or_one(pt)));
  }
  // work space
  const size_t nb = (size_t)c->nbas, ng = (size_t)c->ngrid;
  HIP_TRY(c->lsp.alloc(nb * (2 * nf + 2 * c->nd)));
  c->orb = c->lsp.p;
  c->h01 = c->orb + nb * nf;
  c->gorb = c->h01 + nb * nf;
  c->v2orb = c->gorb + nb * c->nd;
  HIP_TRY(hipMemset(c->lsp.p, 0, c->lsp.n * sizeof(double2)));
  c->mixed = S.precision == 1;
  const size_t np0 = std::max(1, c->npair);
  if (!c->mixed) {
    HIP_TRY(c->gsp.alloc(ng * (nf + 2 * c->nd)));
    c->orbg = c->gsp.p;
    c->gorbg = c->orbg + ng * nf;
    c->v2orbg = c->gorbg + ng * c->nd;
    HIP_TRY(hipMemset(c->gsp.p, 0, c->gsp.n * sizeof(double2)));
    HIP_TRY(c->rhoL.alloc(np0 * L2 * nr));
    HIP_TRY(c->vL.alloc(np0 * L2 * nr));
    HIP_TRY(c->vg.alloc(np0 * nl * nr));
  } else {
    HIP_TRY(c->gsp_f.alloc(ng * (nf + 2 * c->nd)));
    c->orbg_f = c->gsp_f.p;
    c->gorbg_f = c->orbg_f + ng * nf;
    c->v2orbg_f = c->gorbg_f + ng * c->nd;
    HIP_TRY(hipMemset(c->gsp_f.p, 0, c->gsp_f.n * sizeof(float2)));
    HIP_TRY(c->rhoL_f.alloc(np0 * L2 * nr));
    HIP_TRY(c->vL_f.alloc(np0 * L2 * nr));
    HIP_TRY(c->vg_f.alloc(np0 * nl * nr));
    std::vector<double> g2L = padded_cols(S.mat_g2L, S.nlat, S.L2), L2g = padded_cols(S.mat_L2g, S.L2, S.nlat);
    HIP_TRY(c->m_g2L_f.upload(std::vector<float>(g2L.begin(), g2L.end())));
    HIP_TRY(c->m_L2g_f.upload(std::vector<float>(L2g.begin(), L2g.end())));
  }
  HIP_TRY(c->orbg0.alloc(std::max<size_t>(1, ng * nfc)));
  HIP_TRY(hipMemset(c->orbg0.p, 0, c->orbg0.n * sizeof(double2)));
  const size_t np = np0;
  HIP_TRY(c->p_xb.alloc(np * L2 * std::max(1, S.nfe - 1)));
  HIP_TRY(c->p_Q.alloc(np * L2));

  // transform batches
  {
    std::vector<TransformBatch> b;
    for (int f = nfc; f < nf; ++f)
      b.push_back({(long long)(f * nb), 0, (long long)(f * ng), S.mval[f] + S.mmax1, nr});
    c->n_l2g_dyn = (int)b.size();
    HIP_TRY(c->b_l2g_dyn.upload(b));
    b.clear();
    for (int f = 0; f < nfc; ++f)
      b.push_back({(long long)(f * ng), 0, (long long)(f * nb), S.mval[f] + S.mmax1, S.nradfc});
    c->n_g2l_fc = (int)b.size();
    HIP_TRY(c->b_g2l_fc.upload(flat_or_one(b)));
    b.clear();
    // pair densities: in0 = orbital j, in1 = orbital i, matrix for m_i - m_j
    for (int q = 0; q < c->npair; ++q) {
      auto& p = S.pairs[q];
      int mji = S.mval[p.i] - S.mval[p.j];
      b.push_back({(long long)(p.j * ng), (long long)(p.i * ng), (long long)((size_t)q * L2 * nr), mji + S.mmax2,
                   p.type == 1 ? S.nradfc : nr});
    }
    HIP_TRY(c->b_rho.upload(flat_or_one(b)));
    b.clear();
    for (int q = 0; q < c->npair; ++q) {
      auto& p = S.pairs[q];
      int mji = S.mval[p.i] - S.mval[p.j];
      b.push_back({(long long)((size_t)q * L2 * nr), 0, (long long)((size_t)q * nl * nr), mji + S.mmax2,
                   p.type == 1 ? S.nradfc : nr});
    }
    HIP_TRY(c->b_L2g.upload(flat_or_one(b)));
    b.clear();
    // gorbg and v2orbg (contiguous) -> gorb and v2orb (contiguous)
    for (int w = 0; w < 2; ++w)
      for (int f = nfc; f < nf; ++f) {
        int d = w * c->nd + (f - nfc);
        b.push_back({(long long)(d * ng), 0, (long long)(d * nb), S.mval[f] + S.mmax1, nr});
      }
    c->n_g2l_mf = (int)b.size();
    HIP_TRY(c->b_g2l_mf.upload(b));
  }

  // Gram products over the l-space basis: left = orb_0..orb_{nf-1}, h01 of
  // frozen core; right = h01, gorb, v2orb, orb of the dynamical orbitals.
  {
    std::vector<const void*> L, R;
    for (int f = 0; f < nf; ++f) L.push_back(c->orb + (size_t)f * nb);
    for (int f = 0; f < nfc; ++f) L.push_back(c->h01 + (size_t)f * nb);
    for (int f = nfc; f < nf; ++f) R.push_back(c->h01 + (size_t)f * nb);
    for (int d = 0; d < c->nd; ++d) R.push_back(c->gorb + (size_t)d * nb);
    for (int d = 0; d < c->nd; ++d) R.push_back(c->v2orb + (size_t)d * nb);
    for (int f = nfc; f < nf; ++f) R.push_back(c->orb + (size_t)f * nb);
    c->gl_nL = (int)L.size();
    c->gl_nR = (int)R.size();
    if (c->gl_nL * c->gl_nR > kMaxGram) return fail(-2, "too many orbitals for the Gram kernel");
    HIP_TRY(c->g_left_l.upload(L));
    HIP_TRY(c->g_right_l.upload(R));
    std::vector<int> mH(nf * nf, -1), mG(nf * nf, -1), mV(nf * nf, -1), mF(nf * nf, -1);
    const int nR = c->gl_nR, nd = c->nd;
    for (int j = 0; j < nf; ++j)
      for (int i = nfc; i < nf; ++i) {
        if (S.mval[j] != S.mval[i]) continue;
        int di = i - nfc;
        mH[j * nf + i] = 2 * (j * nR + di);
        mG[j * nf + i] = 2 * (j * nR + nd + di);
        mV[j * nf + i] = 2 * (j * nR + 2 * nd + di);
        if (j < nfc) mF[j * nf + i] = 2 * ((nf + j) * nR + 3 * nd + di);
      }
    HIP_TRY(c->mapH.upload(mH));
    HIP_TRY(c->mapG.upload(mG));
    HIP_TRY(c->mapV.upload(mV));
    HIP_TRY(c->mapF.upload(mF));
    HIP_TRY(c->lprod.alloc((size_t)2 * c->gl_nL * c->gl_nR));
  }
  // Gram products over the pair arena for int2e: left = rho of active pairs
  // (i >= j stored as pair (i, j)), right = v of active pairs.
  {
    const int na = S.nact, ncore = S.ncore;
    std::vector<const void*> L, R;
    std::vector<int> act_pairs;
    for (int q = 0; q < c->npair; ++q) {
      auto& p = S.pairs[q];
      if (p.type == 0 && p.i >= ncore) act_pairs.push_back(q);
    }
    for (int q : act_pairs) {
      const size_t off = (size_t)q * L2 * nr;
      if (c->mixed) {
        L.push_back(c->rhoL_f.p + off);
        R.push_back(c->vL_f.p + off);
      } else {
        L.push_back(c->rhoL.p + off);
        R.push_back(c->vL.p + off);
      }
    }
    c->gp_nL = (int)L.size();
    c->gp_nR = (int)R.size();
    if (c->gp_nL * c->gp_nR > kMaxGram) return fail(-2, "too many active pairs for the Gram kernel");
    std::vector<int> m2(std::max(1, na * na * na * na), -1);
    auto slot = [&](int q) { return (int)(std::find(act_pairs.begin(), act_pairs.end(), q) - act_pairs.begin()); };
    for (int ia = 0; ia < na; ++ia)
      for (int ja = 0; ja <= ia; ++ja)
        for (int ka = 0; ka < na; ++ka)
          for (int la = 0; la < na; ++la) {
            int i = ncore + ia, j = ncore + ja, k = ncore + ka, l = ncore + la;
            if (S.mval[i] + S.mval[k] != S.mval[j] + S.mval[l]) continue;
            int qa = S.pair_of[(size_t)i * nf + j];  // stored (i, j), i >= j
            int qb = S.pair_of[(size_t)k * nf + l];
            int conjb = S.pair_conj[(size_t)k * nf + l];
            int e = slot(qa) * c->gp_nR + slot(qb);
            m2[ia + na * (ja + na * (ka + na * la))] = 2 * e + (conjb ? 1 : 0);
          }
    HIP_TRY(c->map2e.upload(m2));
    if (!L.empty()) {
      HIP_TRY(c->g_left_p.upload(L));
      HIP_TRY(c->g_right_p.upload(R));
    }
    HIP_TRY(c->pprod.alloc(std::max<size_t>(1, (size_t)2 * c->gp_nL * c->gp_nR)));
  }
  c->gram_blk = 256;
  c->x_local = S.ncore == S.nfcore || S.nact == 0;
  {
    size_t ent = std::max((size_t)c->gl_nL * c->gl_nR, (size_t)c->gp_nL * c->gp_nR);
    HIP_TRY(c->gram_partial.alloc(std::max<size_t>(1, ent * 2 * c->gram_blk)));
    HIP_TRY(c->gram_partial_p.alloc(std::max<size_t>(1, ent * 2 * c->gram_blk)));
  }
  // island
  const int na = std::max(1, S.nact);
  HIP_TRY(c->den1.alloc(na * na));
  HIP_TRY(c->den2.alloc((size_t)na * na * na * na));
  HIP_TRY(c->rden.alloc(na * na));
  HIP_TRY(c->rrden.alloc(na * na));
  HIP_TRY(c->int1e.alloc(na * na));
  HIP_TRY(c->int2e.alloc((size_t)na * na * na * na));
  HIP_TRY(c->xmat.alloc((size_t)nf * nf));
  HIP_TRY(c->dcic.alloc(S.lcic));
  HIP_TRY(hipMemset(c->den1.p, 0, c->den1.n * sizeof(double2)));
  HIP_TRY(hipMemset(c->den2.p, 0, c->den2.n * sizeof(double2)));
  HIP_TRY(hipMemset(c->rden.p, 0, c->rden.n * sizeof(double2)));
  HIP_TRY(hipMemset(c->rrden.p, 0, c->rrden.n * sizeof(double2)));
  HIP_TRY(c->ene.alloc(5));
  HIP_TRY(hipMemset(c->ene.p, 0, 5 * sizeof(double)));
  const size_t nfull = (size_t)S.nstr_a * S.nstr_b;
  HIP_TRY(c->mfc.alloc((size_t)na * na * na * na + 2 * na * na));
  HIP_TRY(c->dip_part.alloc((size_t)3 * ((S.nr + 255) / 256) * S.L1));  // three forms x dipole_blocks
  HIP_TRY(c->dip3_tmp.alloc(3));
  HIP_TRY(c->ovlp.alloc((size_t)S.nfun * S.nfun));
  HIP_TRY(c->ci_work.alloc(std::max<size_t>(1, ((size_t)2 * S.nact * S.nact + 2) * nfull)));
  HIP_TRY(c->det_of.upload(flat_or_one(S.det_of)));
  HIP_TRY(c->exA.upload(flat_or_one(S.exA)));
  HIP_TRY(c->sgA.upload(flat_or_one(S.sgA)));
  HIP_TRY(c->exB.upload(flat_or_one(S.exB)));
  HIP_TRY(c->sgB.upload(flat_or_one(S.sgB)));
  // RK work
  const size_t wsz = nb * nf + S.lcic;
  HIP_TRY(c->rk_w0.alloc(wsz));
  HIP_TRY(c->rk_t.alloc(wsz));
  HIP_TRY(c->rk_k.alloc(wsz * kMaxStages));
  HIP_TRY(c->lf_dev.alloc(9 * kMaxStages));
  HIP_TRY(c->clock.alloc(1));
  HIP_TRY(side_streams(c->device, c->side));
  // fork/join events order kernels of one device only: device-scope release
  // and acquire suffice, no system-scope fence (cache write-back) per record
  for (auto& e : c->fj) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
  return 0;
}

// With X built in the orbital derivative, the l-space Gram products are left
// as partial sums (no final-sum launch on the main stream): the number of
// partials per entry, else 0.
int l_gram_partials(td1c_ctx* c) {
  if (!c->x_local) return 0;
  GramArgs g{};
  g.len = (size_t)c->nbas;  // the l-space products run over the orbital basis
  return gram_blocks(g, c->gram_blk);
}

IslandArgs island_args(td1c_ctx* c, const double2* cic, double dtime) {
  Setup& S = c->S;
  IslandArgs a{};
  a.nfun = S.nfun; a.nfcore = S.nfcore; a.ncore = S.ncore; a.nact = S.nact; a.lcic = S.lcic;
  a.nelact2 = S.nelact[2]; a.icomp = S.icomp;
  a.mval = c->mval.p;
  a.throcc1 = S.throcc1; a.throcc2 = S.throcc2; a.dtime = dtime;
  a.nstr_a = S.nstr_a; a.nstr_b = S.nstr_b;
  a.det_of = c->det_of.p; a.exA = c->exA.p; a.sgA = c->sgA.p; a.exB = c->exB.p; a.sgB = c->sgB.p;
  a.cic = cic; a.ci_work = c->ci_work.p;
  a.den1 = c->den1.p; a.den2 = c->den2.p; a.rden = c->rden.p; a.rrden = c->rrden.p;
  a.lprod = c->lprod.p;
  a.lpart = c->gram_partial.p;
  a.lnblk = l_gram_partials(c);
  a.mapH = c->mapH.p; a.mapG = c->mapG.p; a.mapV = c->mapV.p; a.mapF = c->mapF.p;
  a.pprod = c->pprod.p; a.map2e = c->map2e.p;
  a.int1e = c->int1e.p; a.int2e = c->int2e.p; a.xmat = c->xmat.p; a.dcic = c->dcic.p;
  a.ene = c->ene.p;
  a.mfc = c->mfc.p;
  island_lds_plan(a);
  return a;
}

// Everything of hprod_htotx up to the X matrix and the CI derivative.
// Runge-Kutta bookkeeping of the CI part of one stage, done by the island
// kernel right after it forms dC (see IslandArgs).
struct CiStage {
  double2* kout = nullptr;
  double2* acc = nullptr;
  double dtw = 0.0;
  double2* next = nullptr;
  const double2* w0 = nullptr;
  double2* w0out = nullptr;
  int nprev = 0;
  const double2* kprev[kMaxStages] = {};
  double cprev[kMaxStages] = {};
  double cnew = 0.0;
};

// hprod_orbin + hprod_fcorb: the current orbitals in the work array (l and
// theta grid), the frozen cores with the velocity-gauge phase of d_lf.
// The frozen-core theta -> l transform goes on stream sfc (after an event
// the caller records on s when sfc differs).
void enqueue_orbin_theta(td1c_ctx* c, const double* d_lf, const double2* wfn, hipStream_t s);
void enqueue_orbin_fc_l(td1c_ctx* c, hipStream_t sfc);

void enqueue_orbin(td1c_ctx* c, const double* d_lf, const double2* wfn, hipStream_t s) {
  enqueue_orbin_theta(c, d_lf, wfn, s);
  enqueue_orbin_fc_l(c, s);
}

void enqueue_orbin_theta(td1c_ctx* c, const double* d_lf, const double2* wfn, hipStream_t s) {
  Setup& S = c->S;
  const int nr = S.nr, nfc = S.nfcore, L1 = S.L1, nl = S.nlat;
  // frozen core in the velocity gauge (hprod_fcorbg): the phase on the theta
  // grid rides on the l -> theta launch of the dynamical orbitals
  const bool fc_vel = nfc > 0 && S.igauge == 1;
  FcPhaseJob fj{};
  fj.orbg0 = c->orbg0.p;
  fj.orbg = c->mixed ? static_cast<void*>(c->orbg_f) : static_cast<void*>(c->orbg);
  fj.xr = c->xr.p; fj.cost = c->cost.p; fj.lf = d_lf;
  fj.nr = nr; fj.nradfc = S.nradfc; fj.nlat = nl; fj.nfc = nfc;
  // dynamical orbitals: copied into the work array by the l -> theta
  // transform that reads them
  {
    TIMED("l2theta");
    if (c->mixed)
      launch_transform_t<double2, float2>(0, wfn, c->orbg_f, c->m_l2g.p, nullptr, c->b_l2g_dyn.p, c->n_l2g_dyn, nr, L1,
                                          nl, pad_cols(nl), c->orb, s, c->tpath, fc_vel ? &fj : nullptr, c->tvar);
    else
      launch_transform_t<double2, double2>(0, wfn, c->orbg, c->m_l2g.p, nullptr, c->b_l2g_dyn.p, c->n_l2g_dyn, nr, L1,
                                           nl, pad_cols(nl), c->orb, s, c->tpath, fc_vel ? &fj : nullptr, c->tvar);
  }
}

void enqueue_orbin_fc_l(td1c_ctx* c, hipStream_t s) {
  Setup& S = c->S;
  const int nr = S.nr, nfc = S.nfcore, L1 = S.L1, nl = S.nlat;
  // the phased frozen core back to l-space
  if (nfc > 0 && S.igauge == 1) {
    TIMED_ON(s, "theta2l_fc");
    if (c->mixed)
      launch_transform_t<float2, double2>(0, c->orbg_f, c->orb, c->m_g2l.p, nullptr, c->b_g2l_fc.p, c->n_g2l_fc, nr, nl,
                                          L1, pad_cols(L1), nullptr, s, c->tpath, nullptr, c->tvar);
    else
      launch_transform(0, c->orbg, c->orb, c->m_g2l.p, c->b_g2l_fc.p, c->n_g2l_fc, nr, nl, L1, pad_cols(L1), nullptr, s,
                       c->tpath, nullptr, c->tvar);
  }
}

// int2e Gram products on side stream s1 once the pair potentials exist on s;
// records fj[4] on s1 when they are done.
hipError_t enqueue_pair_gram(td1c_ctx* c, hipStream_t s, hipStream_t s1) {
  hipError_t e = hipEventRecord(c->fj[0], s);
  if (e == hipSuccess) e = hipStreamWaitEvent(s1, c->fj[0], 0);
  if (e != hipSuccess) return e;
  if (c->gp_nL > 0) {
    GramArgs h{};
    h.len = (size_t)c->S.L2 * c->S.nr; h.nL = c->gp_nL; h.nR = c->gp_nR; h.conj_left = 0; h.with_conj_right = 1;
    h.left = c->g_left_p.p; h.right = c->g_right_p.p; h.f32 = c->mixed ? 1 : 0;
    TIMED_ON(s1, "gram_pair");
    launch_gram(h, c->gram_partial_p.p, c->gram_blk, c->pprod.p, s1);
  }
  return hipEventRecord(c->fj[4], s1);
}

// d_lf: the stage's lfield[9] in device memory.
int enqueue_htotx(td1c_ctx* c, double dtime, const double* d_lf, const double2* wfn, const double2* cic,
                  const CiStage& cs, hipStream_t s) {
  Setup& S = c->S;
  const int nr = S.nr, nf = S.nfun, nfc = S.nfcore, L1 = S.L1, L2 = S.L2, nl = S.nlat;
  const size_t nb = (size_t)c->nbas;
  // Branches: side 0 runs the CI RDM stage and later the pair Gram products,
  // side 1 the one-body part; the main stream runs the mean-field chain.
  hipStream_t s1 = c->side[0], s2 = c->side[1];
  // one fork after the orbitals' l -> theta launch serves both side streams
  // (the RDM stage needs nothing from it, and starts early enough)
  enqueue_orbin_theta(c, d_lf, wfn, s);
  HIP_TRY(hipEventRecord(c->fj[2], s));
  HIP_TRY(hipStreamWaitEvent(s1, c->fj[2], 0));
  IslandArgs ia = island_args(c, cic, dtime);
  if (S.nact > 0) { TIMED_ON(s1, "island_a"); launch_island_a(ia, s1); }
  HIP_TRY(hipEventRecord(c->fj[1], s1));
  HIP_TRY(hipStreamWaitEvent(s2, c->fj[2], 0));
  // side 1: the frozen cores in l-space, then the one-body part (the main
  // stream needs neither before the l-space Gram products)
  enqueue_orbin_fc_l(c, s2);
  // one-body part
  OneBodyArgs ob{};
  ob.nr = nr; ob.L1 = L1; ob.ndvr = S.n; ob.nfcore = nfc; ob.nradfc = S.nradfc;
  ob.mmax1 = S.mmax1; ob.igauge = S.igauge; ob.fc_full = S.iprojfc == 2;
  ob.mval = c->mval.p; ob.band_k = c->band_k.p; ob.band_k0 = c->band_k0.p; ob.band_p = c->band_p.p;
  ob.xr = c->xr.p; ob.alph = c->alph.p; ob.tdiag = c->tdiag.p; ob.lf = d_lf;
  ob.f0 = 0; ob.f1 = nf;
  const OneBodyElArgs oe = onebody_el_args(c, d_lf);
  {
    TIMED_ON(s2, "onebody");
    if (!launch_onebody_el(c->orb, c->h01, oe, s2, c->ob_shape, c->ob_lc)) launch_onebody(c->orb, c->h01, ob, s2);
  }
  HIP_TRY(hipEventRecord(c->fj[3], s2));
  bool pair_gram_done = false;
  bool waited_rdm = false;  // the main stream has joined the RDM stage
  // mean field
  const size_t ng = (size_t)c->ngrid;
  if (S.nel[2] >= 2 && c->npair > 0) {
    if (c->mixed) {
      TIMED("pair_theta2L");
      launch_transform_t<float2, float2>(1, c->orbg_f, c->rhoL_f.p, c->m_g2L.p, c->m_g2L_f.p, c->b_rho.p, c->npair, nr,
                                         nl, L2, pad_cols(L2), nullptr, s, c->tpath, nullptr, c->tvar);
    } else {
      TIMED("pair_theta2L");
      launch_transform(1, c->orbg, c->rhoL.p, c->m_g2L.p, c->b_rho.p, c->npair, nr, nl, L2, pad_cols(L2), nullptr, s, c->tpath, nullptr, c->tvar);
    }
    PoissonArgs pa{};
    pa.nr = nr; pa.n = S.n; pa.nfe = S.nfe; pa.ni = S.ni; pa.L2 = L2; pa.npair = c->npair;
    pa.pair_m = c->pair_m.p;
    pa.G = c->pG.p; pa.gstride = S.nfe; pa.Wl = c->pWl.p; pa.Wr = c->pWr.p; pa.WlI = c->pWlI.p; pa.WrI = c->pWrI.p;
    pa.Sd = c->pSd.p; pa.Sl = c->pSl.p; pa.Su = c->pSu.p;
    pa.pcr_steps = S.pcr_steps; pa.pcrA = c->pcrA.p; pa.pcrG = c->pcrG.p; pa.pcrBinv = c->pcrBinv.p;
    pa.invr = c->invr.p; pa.rpl0 = c->rpl0.p; pa.rpl1 = c->rpl1.p; pa.fac2 = c->fac2.p; pa.fac1 = c->fac1.p;
    pa.Q = c->p_Q.p;
    if (c->mixed) {
      { TIMED("poisson_bridge"); launch_poisson(c->rhoL_f.p, c->p_xb.p, c->vL_f.p, pa, s, 1); }
      { TIMED("poisson_interior"); launch_poisson(c->rhoL_f.p, c->p_xb.p, c->vL_f.p, pa, s, 2); }
      HIP_TRY(enqueue_pair_gram(c, s, s1));
      pair_gram_done = true;
      {
        TIMED("pair_L2theta");
        launch_transform_t<float2, float2>(0, c->vL_f.p, c->vg_f.p, c->m_L2g.p, c->m_L2g_f.p, c->b_L2g.p, c->npair, nr,
                                           L2, nl, pad_cols(nl), nullptr, s, c->tpath, nullptr, c->tvar);
      }
    } else {
      { TIMED("poisson_bridge"); launch_poisson(c->rhoL.p, c->p_xb.p, c->vL.p, pa, s, 1); }
      { TIMED("poisson_interior"); launch_poisson(c->rhoL.p, c->p_xb.p, c->vL.p, pa, s, 2); }
      HIP_TRY(enqueue_pair_gram(c, s, s1));
      pair_gram_done = true;
      { TIMED("pair_L2theta"); launch_transform(0, c->vL.p, c->vg.p, c->m_L2g.p, c->b_L2g.p, c->npair, nr, L2, nl, pad_cols(nl), nullptr, s, c->tpath, nullptr, c->tvar); }
    }
    MfArgs ma{};
    ma.nr = nr; ma.nlat = nl; ma.nfun = nf; ma.nfcore = nfc; ma.nfcore2 = S.nfcore2; ma.ncore = S.ncore; ma.nact = S.nact;
    ma.nradfc = S.nradfc; ma.npair = c->npair; ma.nelact2 = S.nelact[2];
    ma.pj = c->pj.p; ma.pi = c->pi.p; ma.ptype = c->ptype.p; ma.mval = c->mval.p;
    ma.den1 = c->den1.p; ma.den2 = c->den2.p; ma.rden = c->rden.p; ma.mfc = c->mfc.p;
    HIP_TRY(hipStreamWaitEvent(s, c->fj[1], 0));  // RDM stage done
    waited_rdm = true;
    if (c->mixed) {
      { TIMED("mfprod"); launch_mfprod(c->orbg_f, c->vg_f.p, c->gorbg_f, c->v2orbg_f, ma, s); }
      {
        TIMED("theta2l_mf");
        launch_transform_t<float2, double2>(0, c->gorbg_f, c->gorb, c->m_g2l.p, nullptr, c->b_g2l_mf.p, c->n_g2l_mf, nr,
                                            nl, L1, pad_cols(L1), nullptr, s, c->tpath, nullptr, c->tvar);
      }
    } else {
      { TIMED("mfprod"); launch_mfprod(c->orbg, c->vg.p, c->gorbg, c->v2orbg, ma, s); }
      { TIMED("theta2l_mf"); launch_transform(0, c->gorbg, c->gorb, c->m_g2l.p, c->b_g2l_mf.p, c->n_g2l_mf, nr, nl, L1, pad_cols(L1), nullptr, s, c->tpath, nullptr, c->tvar); }
    }
  } else {
    launch_zero(c->gorb, 2 * nb * c->nd, s);
  }
  (void)ng;
  if (!pair_gram_done) HIP_TRY(enqueue_pair_gram(c, s, s1));
  // inner products of the l-space results (needs the one-body branch)
  HIP_TRY(hipStreamWaitEvent(s, c->fj[3], 0));
  // With stage B on side stream 0, the previous hprod's stage B reads lprod
  // until it finishes; it precedes this hprod's RDM stage on that stream, so
  // joining the RDM stage orders the l-space products below after it.
  if (c->x_local && !waited_rdm) HIP_TRY(hipStreamWaitEvent(s, c->fj[1], 0));
  GramArgs g{};
  g.len = nb; g.nL = c->gl_nL; g.nR = c->gl_nR; g.conj_left = 1; g.with_conj_right = 0;
  g.left = c->g_left_l.p; g.right = c->g_right_l.p;
  // (partial sums only when X is built in the orbital derivative: it and
  // stage B sum what they need)
  { TIMED("gram_l"); launch_gram(g, c->gram_partial.p, c->gram_blk, c->x_local ? nullptr : c->lprod.p, s); }
  ia.kout = cs.kout; ia.acc = cs.acc; ia.dtw = cs.dtw; ia.next = cs.next; ia.w0 = cs.w0; ia.w0out = cs.w0out;
  ia.nprev = cs.nprev; ia.cnew = cs.cnew;
  for (int j = 0; j < kMaxStages; ++j) { ia.kprev[j] = cs.kprev[j]; ia.cprev[j] = cs.cprev[j]; }
  if (c->x_local) {
    // the orbital derivative builds X itself: stage B (energies, CI
    // derivative and CI stage update) follows the pair Gram products on side
    // stream 0 once the l-space products exist; callers join it through
    // join_island
    HIP_TRY(hipEventRecord(c->fj[5], s));
    HIP_TRY(hipStreamWaitEvent(s1, c->fj[5], 0));
    { TIMED_ON(s1, "island_b"); launch_island_b(ia, s1); }
    HIP_TRY(hipEventRecord(c->fj[6], s1));
  } else {
    // join the side streams (fj[4] follows fj[1] on side stream 0, so one
    // wait covers the RDM stage and the pair Gram products)
    HIP_TRY(hipStreamWaitEvent(s, c->fj[4], 0));
    { TIMED("island_b"); launch_island_b(ia, s); }
  }
  return 0;
}

DtorbArgs dtorb_args(td1c_ctx* c, double dtime) {
  Setup& S = c->S;
  DtorbArgs d{};
  d.nbas = c->nbas; d.nr = S.nr; d.nfun = S.nfun; d.nfcore = S.nfcore;
  d.mval = c->mval.p; d.h01 = c->h01; d.gorb = c->gorb; d.v2orb = c->v2orb; d.orb = c->orb; d.xmat = c->xmat.p;
  d.tfac = S.icomp == 0 ? make_double2(-dtime, 0.0) : make_double2(0.0, -dtime);
  d.xlocal = c->x_local ? 1 : 0;
  d.ncore = S.ncore;
  d.ladapt = S.lconst;
  d.isplit = S.isplit;
  d.lval = c->lval.p;
  d.xfac = S.icomp == 0 ? make_double2(dtime, 0.0) : make_double2(0.0, dtime);
  d.lprod = c->lprod.p;
  d.mapH = c->mapH.p; d.mapG = c->mapG.p; d.mapV = c->mapV.p; d.mapF = c->mapF.p;
  d.lpart = c->gram_partial.p;
  d.lnblk = l_gram_partials(c);
  return d;
}

// makes the caller's stream wait for the island stage B when it ran on
// side stream 0
hipError_t join_island(td1c_ctx* c, hipStream_t s) {
  return c->x_local ? hipStreamWaitEvent(s, c->fj[6], 0) : hipSuccess;
}

// Host band Cholesky factor of the Poisson operator A_L = 2 K + diag(L(L+1)/r^2)
// (bas_gen_d2ll) for the frozen-core set-up (once per run): lower band
// ab[(n+1)*j + (i-j)].
std::vector<double> host_poisson_factor(const Setup& S, int L) {
  const int nr = S.nr, n = S.n, bw = 2 * n + 1;
  std::vector<double> ab((size_t)(n + 1) * nr, 0.0);
  for (int j = 0; j < nr; ++j)
    for (int i = j; i <= std::min(nr - 1, j + n); ++i) ab[(size_t)(n + 1) * j + (i - j)] = 2.0 * S.radk[(size_t)bw * (j + 1) + n + (i - j)];
  for (int j = 0; j < nr; ++j) ab[(size_t)(n + 1) * j] += (double)(L * (L + 1)) / (S.xr[j] * S.xr[j]);
  for (int j = 0; j < nr; ++j) {
    double d = std::sqrt(ab[(size_t)(n + 1) * j]);
    ab[(size_t)(n + 1) * j] = d;
    int kn = std::min(n, nr - 1 - j);
    for (int i = 1; i <= kn; ++i) ab[(size_t)(n + 1) * j + i] /= d;
    for (int c2 = 1; c2 <= kn; ++c2)
      for (int r2 = c2; r2 <= kn; ++r2)
        ab[(size_t)(n + 1) * (j + c2) + (r2 - c2)] -= ab[(size_t)(n + 1) * j + r2] * ab[(size_t)(n + 1) * j + c2];
  }
  return ab;
}

std::vector<double> host_poisson_solve(const Setup& S, const std::vector<double>& ab, const std::vector<double>& rhs) {
  const int nr = S.nr, n = S.n;
  std::vector<double> x = rhs;
  for (int j = 0; j < nr; ++j) {
    x[j] /= ab[(size_t)(n + 1) * j];
    for (int i = j + 1; i <= std::min(nr - 1, j + n); ++i) x[i] -= x[j] * ab[(size_t)(n + 1) * j + (i - j)];
  }
  for (int j = nr - 1; j >= 0; --j) {
    double t = x[j];
    for (int i = j + 1; i <= std::min(nr - 1, j + n); ++i) t -= ab[(size_t)(n + 1) * j + (i - j)] * x[i];
    x[j] = t / ab[(size_t)(n + 1) * j];
  }
  return x;
}

std::vector<double> host_poisson_l0(const Setup& S, const std::vector<double>& rhs) {
  return host_poisson_solve(S, host_poisson_factor(S, 0), rhs);
}

// Frozen-core energy (hprod_v2fc_enefc, src/hprod/hprod_set_wfn0.F90:790-865),
// once at set-up: 2 <fc|h0|fc> with the bare one-body operator (before
// hprod_addfc_tmat) plus <fc|g|fc>, g = v2jfc (fcj, v2xfc included when
// implicit), v2xfc again on every frozen core when type -2 cores exist (fcx2),
// and minus the exchange with the type -1 cores through the FC-FC pair
// potentials (fcx1; hprod_mkrho2p / mkv2mf_poisson / sph2ang2 / herm 'fcfc'),
// formed on the theta grid and projected back as bas_ang2sph1_fc does.
// orb: frozen cores in l-space [f][l][r]; og: on the theta grid [f][t][r].
double host_fcore_energy(const Setup& S, const std::vector<std::complex<double>>& orb,
                         const std::vector<std::complex<double>>& og, const std::vector<double2>& v2jfc,
                         const double* v2xfc, std::vector<std::complex<double>>* vfc_out) {
  using cd = std::complex<double>;
  const int nr = S.nr, n = S.n, bw = 2 * n + 1, L1 = S.L1, L2 = S.L2, nl = S.nlat, nfc = S.nfcore;
  const int nrf = S.nradfc;
  const size_t nb = (size_t)L1 * nr, ng = (size_t)nl * nr;
  // 2 <fc|h0|fc>
  cd t = 0.0;
  for (int f = 0; f < nfc; ++f)
    for (int l = 0; l < L1; ++l) {
      const std::vector<double>& band = l == 0 ? S.radk : S.radk0;
      const cd* p = orb.data() + (size_t)f * nb + (size_t)l * nr;
      for (int i = 0; i < nrf; ++i) {
        cd h = S.tdiag[2 * ((size_t)l * nr + i)] * p[i];
        for (int j = std::max(0, i - n); j <= std::min(nr - 1, i + n); ++j)
          if (j != i) h += band[(size_t)bw * (i + 1) + n + (j - i)] * p[j];
        t += std::conj(p[i]) * h;
      }
    }
  double e = 2.0 * t.real();
  // g on the theta grid
  std::vector<cd> g(ng * nfc, 0.0);
  for (int f = 0; f < nfc; ++f)
    for (int a = 0; a < nl; ++a)
      for (int i = 0; i < nrf; ++i) {
        cd v(v2jfc[i].x, v2jfc[i].y);
        if (S.nfcore2 > 0 && v2xfc) v += cd(v2xfc[2 * i], v2xfc[2 * i + 1]);
        g[(size_t)f * ng + (size_t)a * nr + i] += v * og[(size_t)f * ng + (size_t)a * nr + i];
      }
  if (S.nfcore1 > 0) {
    // pair potentials v(a, b) of the pairs a >= b on the theta grid
    std::vector<cd> vg((size_t)nfc * nfc * ng, 0.0), rl((size_t)L2 * nr), vl((size_t)L2 * nr);
    std::vector<double> re(nr), im(nr);
    std::vector<std::vector<double>> fac(L2);  // band factors, made on first use
    for (int b = 0; b < nfc; ++b)
      for (int a = b; a < nfc; ++a) {
        const int mji = -S.mval[a] + S.mval[b], am = std::abs(mji);
        const cd* oa = og.data() + (size_t)a * ng;
        const cd* ob = og.data() + (size_t)b * ng;
        std::fill(rl.begin(), rl.end(), 0.0);
        for (int L = am; L < L2; ++L)
          for (int q = 0; q < nl; ++q) {
            const double w = S.mat_g2L[((size_t)(mji + S.mmax2) * nl + q) * L2 + L];
            for (int i = 0; i < nrf; ++i) rl[(size_t)L * nr + i] += std::conj(oa[(size_t)q * nr + i]) * ob[(size_t)q * nr + i] * w;
          }
        std::fill(vl.begin(), vl.end(), 0.0);
        for (int L = am; L < L2; ++L) {
          const double* rpl0 = &S.d2rpl0[(size_t)L * nr];
          const double* invr = &S.d2invr[(size_t)L * nr];
          double tr = 0.0, ti = 0.0;
          for (int i = 0; i < nr; ++i) {
            const cd r = rl[(size_t)L * nr + i];
            tr += r.real() * rpl0[i];
            ti += r.imag() * rpl0[i];
            re[i] = r.real() * invr[i];
            im[i] = r.imag() * invr[i];
          }
          tr *= S.d2fac1[L];
          ti *= S.d2fac1[L];
          if (fac[L].empty()) fac[L] = host_poisson_factor(S, L);
          const std::vector<double> xr = host_poisson_solve(S, fac[L], re), xi = host_poisson_solve(S, fac[L], im);
          for (int i = 0; i < nr; ++i) {
            const double f2 = S.d2fac2[(size_t)L * nr + i], r1 = S.d2rpl1[(size_t)L * nr + i];
            vl[(size_t)L * nr + i] = cd((xr[i] + tr * r1) * f2, (xi[i] + ti * r1) * f2);
          }
        }
        cd* v = vg.data() + ((size_t)a * nfc + b) * ng;
        for (int q = 0; q < nl; ++q)
          for (int L = am; L < L2; ++L) {
            const double w = S.mat_L2g[((size_t)(mji + S.mmax2) * L2 + L) * nl + q];
            for (int i = 0; i < nr; ++i) v[(size_t)q * nr + i] += vl[(size_t)L * nr + i] * w;
          }
      }
    if (vfc_out) {
      // every v(j, i) as vfc[(i nfc + j) ng + x] (the potential of conj(phi_j) phi_i)
      vfc_out->assign((size_t)nfc * nfc * ng, 0.0);
      for (int i = 0; i < nfc; ++i)
        for (int j = 0; j < nfc; ++j)
          for (size_t x = 0; x < ng; ++x)
            (*vfc_out)[((size_t)i * nfc + j) * ng + x] =
                j >= i ? vg[((size_t)j * nfc + i) * ng + x] : std::conj(vg[((size_t)i * nfc + j) * ng + x]);
    }
    // fcx1: g_i -= v(j, i) phi_j for the type -1 cores j; v(j, i) = conj(v(i, j)) for j < i
    for (int i = 0; i < nfc; ++i)
      for (int j = S.nfcore2; j < nfc; ++j)
        for (int q = 0; q < nl; ++q)
          for (int r = 0; r < nrf; ++r) {
            const size_t x = (size_t)q * nr + r;
            const cd v = j >= i ? vg[((size_t)j * nfc + i) * ng + x] : std::conj(vg[((size_t)i * nfc + j) * ng + x]);
            g[(size_t)i * ng + x] -= v * og[(size_t)j * ng + x];
          }
  }
  // back to l-space (rows < nradfc) and <fc|g|fc>
  t = 0.0;
  for (int f = 0; f < nfc; ++f) {
    const int m = S.mval[f];
    for (int l = 0; l < L1; ++l)
      for (int i = 0; i < nrf; ++i) {
        cd s = 0.0;
        for (int q = 0; q < nl; ++q)
          s += S.mat_g2l[((size_t)(m + S.mmax1) * nl + q) * L1 + l] * g[(size_t)f * ng + (size_t)q * nr + i];
        t += std::conj(orb[(size_t)f * nb + (size_t)l * nr + i]) * s;
      }
  }
  return e + t.real();
}

constexpr double kPi = 3.14159265358979323846264338327950288;

}  // namespace

extern "C" {

const char* td1c_last_error(void) { return g_error.c_str(); }

// ---------------------------------------------------------------------------
// Autotuner of launch shapes.  Candidates of a kernel are timed on the
// context's own buffers (HIP events, a private stream, one warm-up and five
// timed launches each) and the fastest is kept; every candidate computes the
// same bits, so the choice changes speed only.  Results are cached per
// process by problem shape, so contexts of the same shape tune once.
// TD1C_AUTOTUNE=0 keeps the built-in defaults.
// ---------------------------------------------------------------------------
extern "C++" {
namespace {

std::string onebody_shape_key(const Setup& S) {
  return std::to_string(S.nr) + "/" + std::to_string(S.nfe) + "/" + std::to_string(S.L1) + "/" +
         std::to_string(S.nfun) + "/" + std::to_string(S.nfcore) + "/" + std::to_string(S.igauge);
}

int autotune_onebody(td1c_ctx* c) {
  static std::mutex mu;
  static std::map<std::string, std::pair<int, int>> cache;
  const Setup& S = c->S;
  if (S.n != 20) return 0;  // the band kernel runs for other ndvr
  const char* env = std::getenv("TD1C_AUTOTUNE");
  if (env && std::string(env) == "0") return 0;
  const std::string key = onebody_shape_key(S);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      c->ob_shape = it->second.first;
      c->ob_lc = it->second.second;
      return 0;
    }
  }
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const OneBodyElArgs oe = onebody_el_args(c, c->lf_dev.p);
  HIP_TRY(hipMemsetAsync(c->orb, 0, (size_t)c->nbas * S.nfun * sizeof(double2), st));
  HIP_TRY(hipMemsetAsync(c->lf_dev.p, 0, 9 * sizeof(double), st));
  float best = 1e30f;
  int bs = c->ob_shape, bl = c->ob_lc;
  std::string log = "{\"onebody\": [";
  const int lcs[] = {1, 2, 4, 8, 16};
  bool first = true;
  for (int shape = 0; shape < 3; ++shape)
    for (int lc : lcs) {
      if (lc > S.L1) continue;
      launch_onebody_el(c->orb, c->h01, oe, st, shape, lc);
      HIP_TRY(hipEventRecord(e0, st));
      for (int r = 0; r < 5; ++r) launch_onebody_el(c->orb, c->h01, oe, st, shape, lc);
      HIP_TRY(hipEventRecord(e1, st));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      ms /= 5.0f;
      char buf[96];
      std::snprintf(buf, sizeof buf, "%s[%d, %d, %.2f]", first ? "" : ", ", shape, lc, ms * 1e3f);
      log += buf;
      first = false;
      if (ms < best) {
        best = ms;
        bs = shape;
        bl = lc;
      }
    }
  log += "]}";
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  HIP_TRY(hipGetLastError());
  c->ob_shape = bs;
  c->ob_lc = bl;
  c->tune_log = log;
  std::lock_guard<std::mutex> lock(mu);
  cache[key] = {bs, bl};
  return 0;
}

}  // namespace
}  // extern "C++"

int td1c_ctx_create(const td1c_problem* prob, td1c_ctx** out) {
  if (!prob || !out) return fail(-1, "null argument");
  std::unique_ptr<td1c_ctx> c(new td1c_ctx());
  if (!c->S.build(*prob)) return fail(-2, c->S.error);
  if (c->S.nfun > kMaxFun) return fail(-2, "more than 16 orbitals are not provided by the mean-field kernel");
  if (c->S.nact > kMaxAct) return fail(-2, "more than 8 active orbitals are not provided by the mean-field kernel");
  if (c->S.L2 > 256 || c->S.nlat > 256 || c->S.L1 > 256) return fail(-2, "lmax1/lmax2/nlat above 255 are not provided");
  if (c->S.ni > 31) return fail(-2, "ndvr above 32 is not provided by the Poisson kernel");
  if (c->S.nfe - 1 > kMaxBridge) return fail(-2, "more than 161 finite elements are not provided by the Poisson kernel");
  int rc = build_device(c.get());
  if (rc) return rc;
  rc = autotune_onebody(c.get());
  if (rc) return rc;
  if (const char* tv = std::getenv("TD1C_TVAR")) c->tvar = std::max(0, std::min(5, std::atoi(tv)));
  *out = c.release();
  return 0;
}

int td1c_set_onebody_shape(td1c_ctx* c, int elements_per_block, int l_per_block) {
  if (!c) return fail(-1, "null argument");
  const int shape = elements_per_block == 3 ? 0 : elements_per_block == 6 ? 1 : elements_per_block == 12 ? 2 : -1;
  if (shape < 0 || l_per_block < 1 || l_per_block > 64)
    return fail(-1, "elements_per_block must be 3, 6 or 12 and l_per_block 1..64");
  if (c->gexec) {  // a captured step holds the old launch shape
    HIP_TRY(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  c->ob_shape = shape;
  c->ob_lc = l_per_block;
  return 0;
}

int td1c_get_shapes(const td1c_ctx* c, char* buf, int buflen) {
  if (!c || !buf || buflen <= 0) return fail(-1, "null argument");
  std::string js = "{\"onebody_elements_per_block\": " + std::to_string(c->ob_shape == 0 ? 3 : c->ob_shape == 1 ? 6 : 12) +
                   ", \"onebody_l_per_block\": " + std::to_string(c->ob_lc) +
                   ", \"measured_us\": " + (c->tune_log.empty() ? std::string("null") : c->tune_log) + "}";
  if ((int)js.size() + 1 > buflen) return fail(-1, "buffer too small");
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return 0;
}

int td1c_ctx_destroy(td1c_ctx* ctx) {
  if (!ctx) return 0;
  if (g_bound == ctx) g_bound = nullptr;
  // the side streams outlive the context: let its work on them finish
  // before the buffers go
  for (auto& q : ctx->side)
    if (q) (void)hipStreamSynchronize(q);
  delete ctx;
  return 0;
}

// The orbital counts and CAS string tables of a problem, without its grid.
static int ci_setup(const td1c_problem* p, Setup& S) {
  S.nfcore = p->nfcore2 + p->nfcore1;
  S.ncore = S.nfcore + p->ndcore;
  S.nact = p->nact;
  S.nfun = S.ncore + S.nact;
  if (S.nact < 0 || S.nact > 16 || S.ncore < 0) return fail(-1, "bad orbital counts (nact <= 16)");
  S.mval.assign(S.nfun, 0);
  for (int i = 0; i < S.nfun; ++i) S.mval[i] = p->mval ? p->mval[i] : 0;
  S.nelact[0] = p->nel_alpha - S.ncore;
  S.nelact[1] = p->nel_beta - S.ncore;
  if (S.nelact[0] < 0 || S.nelact[1] < 0 || S.nelact[0] > S.nact || S.nelact[1] > S.nact)
    return fail(-1, "electron count does not fit the orbital classes");
  S.mtot = p->mtot;
  if (!S.build_ci()) return fail(-1, S.error);
  return 0;
}

int td1c_ci_layout(const td1c_problem* p, int* nstr_alpha, int* nstr_beta, int* lcic, int* cic0,
                   unsigned int* str_alpha, unsigned int* str_beta, int* det_alpha, int* det_beta) {
  if (!p) return fail(-1, "null argument");
  Setup S;
  if (int rc = ci_setup(p, S)) return rc;
  if (nstr_alpha) *nstr_alpha = S.nact > 0 ? S.nstr_a : 0;
  if (nstr_beta) *nstr_beta = S.nact > 0 ? S.nstr_b : 0;
  if (lcic) *lcic = S.lcic;
  if (cic0) *cic0 = S.nact > 0 ? S.cic0 : -1;
  if (S.nact > 0) {
    if (str_alpha) std::copy(S.str_a.begin(), S.str_a.end(), str_alpha);
    if (str_beta) std::copy(S.str_b.begin(), S.str_b.end(), str_beta);
    for (int ib = 0; ib < S.nstr_b; ++ib)
      for (int ia = 0; ia < S.nstr_a; ++ia) {
        const int d = S.det_of[(size_t)ib * S.nstr_a + ia];
        if (d < 0) continue;
        if (det_alpha) det_alpha[d] = ia;
        if (det_beta) det_beta[d] = ib;
      }
  } else {
    if (det_alpha) det_alpha[0] = 0;
    if (det_beta) det_beta[0] = 0;
  }
  return 0;
}

int td1c_get_sizes(const td1c_ctx* c, td1c_sizes* out) {
  if (!c || !out) return fail(-1, "null argument");
  out->nrad = c->S.nrad;
  out->nbas = c->nbas;
  out->nfun = c->S.nfun;
  out->lcic = c->S.lcic;
  out->nlat = c->S.nlat;
  out->lmax1 = c->S.lmax1;
  out->lmax2 = c->S.lmax2;
  out->npair = c->npair;
  out->wfn_size = (long long)c->nbas * c->S.nfun + c->S.lcic;
  return 0;
}

int td1c_set_wfn0(td1c_ctx* c, const double* wfn, const double* cic, const double* v2xfc) {
  if (!c || !wfn) return fail(-1, "null argument");
  (void)cic;
  Setup& S = c->S;
  const int nr = S.nr, nfc = S.nfcore, L1 = S.L1;
  const size_t nb = (size_t)c->nbas;
  c->wfn0_set = true;
  if (nfc == 0) return 0;
  const std::complex<double>* w = reinterpret_cast<const std::complex<double>*>(wfn);
  // frozen-core orbitals (orb0) and their theta-grid image at A = 0
  HIP_TRY(hipMemcpy(c->orb, wfn, nb * nfc * sizeof(double2), hipMemcpyHostToDevice));
  {
    std::vector<TransformBatch> b;
    for (int f = 0; f < nfc; ++f)
      b.push_back({(long long)(f * nb), 0, (long long)(f * (size_t)c->ngrid), S.mval[f] + S.mmax1, S.nradfc});
    DevBuf<TransformBatch> tb;
    HIP_TRY(tb.upload(b));
    launch_transform(0, c->orb, c->orbg0.p, c->m_l2g.p, tb.p, nfc, nr, L1, S.nlat, pad_cols(S.nlat), nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    tb.release();
    if (c->mixed) {
      launch_to_f32(c->orbg_f, c->orbg0.p, (size_t)c->ngrid * nfc, nullptr);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipDeviceSynchronize());
    } else {
      HIP_TRY(hipMemcpy(c->orbg, c->orbg0.p, (size_t)c->ngrid * nfc * sizeof(double2), hipMemcpyDeviceToDevice));
    }
  }
  // frozen-core orbitals as hprod_orbin sees them at lfield = 0
  std::vector<std::complex<double>> orbfc(nb * nfc);
  if (S.igauge == 0) {
    std::copy(w, w + nb * nfc, orbfc.begin());
  } else {
    // l -> theta -> l round trip, as hprod_fcorbg + bas_ang2sph1_fc do at A = 0
    DevBuf<TransformBatch> tb;
    std::vector<TransformBatch> b;
    for (int f = 0; f < nfc; ++f)
      b.push_back({(long long)(f * (size_t)c->ngrid), 0, (long long)(f * nb), S.mval[f] + S.mmax1, S.nradfc});
    HIP_TRY(tb.upload(b));
    DevBuf<double2> tmp;
    HIP_TRY(tmp.alloc(nb * nfc));
    launch_transform(0, c->orbg0.p, tmp.p, c->m_g2l.p, tb.p, nfc, nr, S.nlat, L1, pad_cols(L1), nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(orbfc.data(), tmp.p, nb * nfc * sizeof(double2), hipMemcpyDeviceToHost));
    tmp.release();
    tb.release();
  }
  // frozen-core density and Hartree potential (hprod_mkrhofc, hprod_mkv2jfc,
  // src/hprod/hprod_set_wfn0.F90:221-385)
  std::vector<double> rho(nr, 0.0);
  for (int f = 0; f < nfc; ++f) {
    if (S.mval[f] != 0) continue;
    double lfac = 2.0 * S.lval[f] + 1.0;
    for (int i = 0; i < S.nradfc; ++i) {
      std::complex<double> z = orbfc[(size_t)f * nb + (size_t)S.lval[f] * nr + i];
      rho[i] += std::norm(z) * lfac;
    }
  }
  double tr = 0.0;
  std::vector<double> rhs(nr);
  for (int i = 0; i < nr; ++i) {
    tr += rho[i] * S.d2rpl0[i];
    rhs[i] = rho[i] * S.d2invr[i];
  }
  tr *= 1.0 / S.xrad[nr + 1];
  std::vector<double> x = host_poisson_l0(S, rhs);
  std::vector<double2> v2jfc(nr);
  for (int i = 0; i < nr; ++i) v2jfc[i] = make_double2((x[i] + tr * S.d2rpl1[i]) * S.d2fac2[i], 0.0);
  if (S.nfcore2 > 0 && v2xfc)
    for (int i = 0; i < nr; ++i) {
      v2jfc[i].x += v2xfc[2 * i];
      v2jfc[i].y += v2xfc[2 * i + 1];
    }
  {
    std::vector<std::complex<double>> og((size_t)c->ngrid * nfc);
    HIP_TRY(hipMemcpy(og.data(), c->orbg0.p, og.size() * sizeof(double2), hipMemcpyDeviceToHost));
    std::vector<std::complex<double>> vfc;
    c->ene_fcore = host_fcore_energy(S, orbfc, og, v2jfc, v2xfc, &vfc);
    // the frozen-core dipole corrections' tables and work arrays
    const size_t ng = (size_t)c->ngrid;
    if (vfc.empty()) vfc.assign((size_t)nfc * nfc * ng, 0.0);
    c->vfc_tab.release();
    HIP_TRY(c->vfc_tab.alloc(vfc.size()));
    HIP_TRY(hipMemcpy(c->vfc_tab.p, vfc.data(), vfc.size() * sizeof(double2), hipMemcpyHostToDevice));
    if (!c->gfc_l.p) {
      HIP_TRY(c->gfc_l.alloc((size_t)2 * nfc * nb));
      HIP_TRY(hipMemset(c->gfc_l.p, 0, (size_t)2 * nfc * nb * sizeof(double2)));
      if (c->mixed) HIP_TRY(c->gfc_gf.alloc((size_t)2 * nfc * ng));
      else HIP_TRY(c->gfc_g.alloc((size_t)2 * nfc * ng));
      std::vector<TransformBatch> b;
      for (int k = 0; k < 2 * nfc; ++k)
        b.push_back({(long long)(k * ng), 0, (long long)(k * nb), S.mval[k % nfc] + S.mmax1, S.nradfc});
      HIP_TRY(c->b_g2l_gfc.upload(b));
      DipoleFcArgs probe{};
      probe.nfun = S.nfun; probe.nfcore = nfc; probe.nradfc = S.nradfc; probe.L1 = S.L1;
      HIP_TRY(c->dipfc_part.alloc((size_t)dipole_fc_quantities(probe) * dipole_fc_blocks(probe)));
    }
  }
  // the Crank-Nicolson factors of split_type = 1 hold tmat: rebuilt with v2jfc on next use
  c->v2jfc_h.assign((size_t)2 * nr, 0.0);
  for (int i = 0; i < nr; ++i) {
    c->v2jfc_h[2 * i] = v2jfc[i].x;
    c->v2jfc_h[2 * i + 1] = v2jfc[i].y;
  }
  c->cn_dt = 0.0;
  // tmat += v2jfc (hprod_addfc_tmat)
  std::vector<double2> td((size_t)L1 * nr), tz((size_t)L1 * nr);
  for (int l = 0; l < L1; ++l)
    for (int i = 0; i < nr; ++i) {
      size_t q = (size_t)l * nr + i;
      td[q] = make_double2(S.tdiag[2 * q] + v2jfc[i].x, S.tdiag[2 * q + 1] + v2jfc[i].y);
      tz[q] = make_double2(S.tzor[2 * q] + v2jfc[i].x, S.tzor[2 * q + 1] + v2jfc[i].y);
    }
  HIP_TRY(hipMemcpy(c->tdiag.p, td.data(), td.size() * sizeof(double2), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->tzor.p, tz.data(), tz.size() * sizeof(double2), hipMemcpyHostToDevice));
  return 0;
}

int td1c_htot(td1c_ctx* c, double dtime, const double* lfield, const double* d_wfn, const double* d_cic,
              double* d_hwfn, double* d_hcic, void* stream) {
  if (!c || !lfield || !d_wfn || !d_hwfn) return fail(-1, "null argument");
  if (c->S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  hipStream_t s = (hipStream_t)stream;
  const double2* wfn = reinterpret_cast<const double2*>(d_wfn);
  const double2* cic = reinterpret_cast<const double2*>(d_cic);
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  CiStage cs;
  cs.kout = reinterpret_cast<double2*>(d_hcic);
  int rc = enqueue_htotx(c, dtime, c->lf_dev.p, wfn, cic, cs, s);
  if (rc) return rc;
  DtorbArgs d = dtorb_args(c, dtime);
  d.kout = reinterpret_cast<double2*>(d_hwfn);
  { TIMED("dtorb"); launch_dtorb(d, s); }
  HIP_TRY(join_island(c, s));
  HIP_TRY(hipGetLastError());
  return 0;
}

// hprod_htotx (src/hprod/hprod_htot.F90:44-208; clhprod::htotx,
// src/hprod/clhprod.cpp:84-88): everything of the hprod but the orbital
// derivative (dtorb): RDMs, integrals, X matrix, energies and the CI
// derivative, left in the context for the getters below.
int td1c_htotx(td1c_ctx* c, double dtime, const double* lfield, const double* d_wfn, const double* d_cic,
               void* stream) {
  if (!c || !lfield || !d_wfn) return fail(-1, "null argument");
  if (c->S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (c->S.nact > 0 && !d_cic) return fail(-1, "null CI vector");
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  CiStage cs;
  int rc = enqueue_htotx(c, dtime, c->lf_dev.p, reinterpret_cast<const double2*>(d_wfn),
                         reinterpret_cast<const double2*>(d_cic), cs, s);
  if (rc) return rc;
  HIP_TRY(join_island(c, s));
  HIP_TRY(hipGetLastError());
  return 0;
}

// The module arrays of the last td1c_htot / td1c_htotx, as hprod_getden1,
// hprod_getint1, hprod_getint2 (src/hprod/hprod_getvar.F90:2-36) return them:
// column-major (Fortran) complex128 into host memory; synchronises the device.
static int copy_result(const DevBuf<double2>& b, size_t n, double* out) {
  HIP_TRY(hipDeviceSynchronize());
  if (n > 0) HIP_TRY(hipMemcpy(out, b.p, n * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}
int td1c_get_den1(td1c_ctx* c, double* out) {
  if (!c || !out) return fail(-1, "null argument");
  const size_t n = (size_t)c->S.nact;
  return copy_result(c->den1, n * n, out);
}
int td1c_get_den2(td1c_ctx* c, double* out) {
  if (!c || !out) return fail(-1, "null argument");
  const size_t n = (size_t)c->S.nact;
  return copy_result(c->den2, n * n * n * n, out);
}
int td1c_get_int1e(td1c_ctx* c, double* out) {
  if (!c || !out) return fail(-1, "null argument");
  const size_t n = (size_t)c->S.nact;
  return copy_result(c->int1e, n * n, out);
}
int td1c_get_int2e(td1c_ctx* c, double* out) {
  if (!c || !out) return fail(-1, "null argument");
  const size_t n = (size_t)c->S.nact;
  return copy_result(c->int2e, n * n * n * n, out);
}
int td1c_get_xmat(td1c_ctx* c, double* out) {
  if (!c || !out) return fail(-1, "null argument");
  const size_t n = (size_t)c->S.nfun;
  return copy_result(c->xmat, n * n, out);
}

int td1c_get_energies(td1c_ctx* c, double ene[5]) {
  if (!c || !ene) return fail(-1, "null argument");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(ene, c->ene.p, 5 * sizeof(double), hipMemcpyDeviceToHost));
  ene[0] = c->ene_fcore;
  ene[2] += c->ene_fcore;
  ene[4] += c->ene_fcore;
  return 0;
}

// orbitals (hprod_orbin) and 1-RDM of a wave function, then the dipole traces
static int enqueue_dipole(td1c_ctx* c, int nforms, const double* lfield, const double* d_wfn, const double* d_cic,
                          double* d_out, void* stream) {
  if (!c || !lfield || !d_wfn || !d_out) return fail(-1, "null argument");
  if (c->S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (c->S.nact > 0 && !d_cic) return fail(-1, "null CI vector");
  hipStream_t s = (hipStream_t)stream;
  Setup& S = c->S;
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  enqueue_orbin(c, c->lf_dev.p, reinterpret_cast<const double2*>(d_wfn), s);
  IslandArgs ia = island_args(c, reinterpret_cast<const double2*>(d_cic), 1.0);
  if (S.nact > 0) launch_island_a(ia, s);
  DipoleArgs a{};
  a.nr = S.nr; a.L1 = S.L1; a.nfun = S.nfun; a.nfcore = S.nfcore; a.ncore = S.ncore; a.nact = S.nact;
  a.nradfc = S.nradfc; a.mmax1 = S.mmax1; a.mval = c->mval.p; a.xr = c->xr.p; a.alph = c->alph.p;
  a.orb = c->orb; a.den1 = c->den1.p;
  a.ndvr = S.n; a.igauge = S.igauge; a.znuc = S.znuc; a.nel = (double)S.nel[2];
  a.band_p = c->band_p.p; a.lf = c->lf_dev.p;
  launch_dipole(a, nforms, c->dip_part.p, d_out, s);
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_dipole(td1c_ctx* c, const double* lfield, const double* d_wfn, const double* d_cic, double* d_out,
                void* stream) {
  return enqueue_dipole(c, 1, lfield, d_wfn, d_cic, d_out, stream);
}

int td1c_dipole3(td1c_ctx* c, const double* lfield, const double* d_wfn, const double* d_cic, double* d_out,
                 void* stream) {
  return enqueue_dipole(c, 3, lfield, d_wfn, d_cic, d_out, stream);
}

// hprod_dipole (src/hprod/hprod_dipole.F90:2-146) with the frozen-core
// corrections: out9 = {dip, vel, acc} x {(1), (2), (3)}.  With frozen cores
// the generalised Fock vector (hprod_gfock) is one hprod's intermediate
// arrays (the one-body part of every orbital, the mean field of the
// dynamical orbitals) plus the frozen cores' own mean-field rows
// (k_gfock_fc, then theta -> l); k_dipole_fc forms the correction sums.
int td1c_dipole_all(td1c_ctx* c, const double* lfield, const double* d_wfn, const double* d_cic, double* d_out,
                    void* stream) {
  if (!c || !lfield || !d_wfn || !d_out) return fail(-1, "null argument");
  Setup& S = c->S;
  hipStream_t s = (hipStream_t)stream;
  if (S.nfcore == 0) {
    if (int rc = enqueue_dipole(c, 3, lfield, d_wfn, d_cic, c->dip3_tmp.p, stream)) return rc;
    launch_dipole_expand(c->dip3_tmp.p, d_out, s);
    HIP_TRY(hipGetLastError());
    return 0;
  }
  if (!c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (S.nact > 0 && !d_cic) return fail(-1, "null CI vector");
  const double2* wfn = reinterpret_cast<const double2*>(d_wfn);
  const double2* cic = reinterpret_cast<const double2*>(d_cic);
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  CiStage cs;
  if (int rc = enqueue_htotx(c, 1.0, c->lf_dev.p, wfn, cic, cs, s)) return rc;
  HIP_TRY(join_island(c, s));
  const int nfc = S.nfcore, nr = S.nr, nl = S.nlat, L1 = S.L1;
  GfockFcArgs ga{};
  ga.nr = nr; ga.nlat = nl; ga.nradfc = S.nradfc; ga.nfun = S.nfun; ga.nfcore = nfc; ga.nfcore2 = S.nfcore2;
  ga.ncore = S.ncore; ga.nact = S.nact; ga.npair = c->npair;
  ga.pj = c->pj.p; ga.pi = c->pi.p; ga.ptype = c->ptype.p; ga.mval = c->mval.p; ga.den1 = c->den1.p;
  ga.vfc = c->vfc_tab.p;
  if (c->mixed) {
    launch_gfock_fc<float2>(c->orbg_f, c->vg_f.p, c->gfc_gf.p, ga, s);
    launch_transform_t<float2, double2>(0, c->gfc_gf.p, c->gfc_l.p, c->m_g2l.p, nullptr, c->b_g2l_gfc.p, 2 * nfc, nr,
                                        nl, L1, pad_cols(L1), nullptr, s, c->tpath, nullptr, c->tvar);
  } else {
    launch_gfock_fc<double2>(c->orbg, c->vg.p, c->gfc_g.p, ga, s);
    launch_transform(0, c->gfc_g.p, c->gfc_l.p, c->m_g2l.p, c->b_g2l_gfc.p, 2 * nfc, nr, nl, L1, pad_cols(L1),
                     nullptr, s, c->tpath, nullptr, c->tvar);
  }
  DipoleArgs da{};
  da.nr = nr; da.L1 = L1; da.nfun = S.nfun; da.nfcore = nfc; da.ncore = S.ncore; da.nact = S.nact;
  da.nradfc = S.nradfc; da.mmax1 = S.mmax1; da.mval = c->mval.p; da.xr = c->xr.p; da.alph = c->alph.p;
  da.orb = c->orb; da.den1 = c->den1.p;
  da.ndvr = S.n; da.igauge = S.igauge; da.znuc = S.znuc; da.nel = (double)S.nel[2];
  da.band_p = c->band_p.p; da.lf = c->lf_dev.p;
  launch_dipole(da, 3, c->dip_part.p, c->dip3_tmp.p, s);
  DipoleFcArgs fa{};
  fa.nr = nr; fa.L1 = L1; fa.nfun = S.nfun; fa.nfcore = nfc; fa.ncore = S.ncore; fa.nact = S.nact;
  fa.nradfc = S.nradfc; fa.mmax1 = S.mmax1; fa.ndvr = S.n; fa.igauge = S.igauge;
  fa.mval = c->mval.p; fa.xr = c->xr.p; fa.alph = c->alph.p; fa.band_p = c->band_p.p; fa.lf = c->lf_dev.p;
  const size_t nb = (size_t)c->nbas;
  fa.orb = c->orb; fa.h01 = c->h01; fa.gdyn = c->gorb; fa.vdyn = c->v2orb;
  fa.gfc = c->gfc_l.p; fa.vfc = c->gfc_l.p + (size_t)nfc * nb;
  fa.den1 = c->den1.p;
  const size_t need = (size_t)dipole_fc_quantities(fa) * dipole_fc_blocks(fa);
  if (c->dipfc_part.n < need) return fail(-2, "dipole correction buffer too small");
  launch_dipole_fc(fa, c->dipfc_part.p, c->dip3_tmp.p, d_out, s);
  HIP_TRY(hipGetLastError());
  return 0;
}

// Ionisation probabilities (hprod_ipx, src/hprod/hprod_ipx.F90:2-19, with
// ormas_ipx_core / ormas_ipx_ras, src/ormas/ormas_ipx_core.F90:2-130,
// src/ormas/ormas_ipx_ras.F90:2-180): ipx[k] = probability that exactly k
// electrons are outside rad_ipx, k = 0..max_ipx.
//
// The windowed overlap S_in (hprod_mkovlp: rows up to and including the
// first r > rad_ipx) of the orbitals of hprod_orbin is formed on the device.
// The counting statistics follow from the generating function
//   F(x) = <Psi| prod_electrons (P_in + x P_out) |Psi>
//        = sum_IJ conj(C_I) C_J det A^a_IJ(x) det A^b_IJ(x),
//   A_IJ(x)[e][f] = (1 - x) S_in(bra_e, ket_f) + x delta(bra_e, ket_f),
// a polynomial of degree N whose coefficient k is ipx[k]: F is evaluated at
// the N+1 roots of unity and transformed back.  The reference reaches the
// same coefficients through sums of minors and a binomial inversion.
extern "C++" {
namespace {

std::complex<double> host_det(int n, std::vector<std::complex<double>>& a) {
  std::complex<double> d = 1.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(a[(size_t)i * n + k]) > std::abs(a[(size_t)p * n + k])) p = i;
    if (a[(size_t)p * n + k] == 0.0) return 0.0;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
      d = -d;
    }
    d *= a[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const std::complex<double> f = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      for (int j = k; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
    }
  }
  return d;
}

// det A_IJ(x) for every string pair of one spin: out[I * ns + J]
void spin_dets(const Setup& S, const std::vector<std::complex<double>>& ov, const std::vector<uint32_t>& str, int nel,
               std::complex<double> x, std::vector<std::complex<double>>& out) {
  const int nf = S.nfun, nc = S.ncore, ns = (int)str.size();
  out.assign((size_t)ns * ns, 0.0);
  std::vector<int> bra(nel), ket(nel);
  std::vector<std::complex<double>> a((size_t)nel * nel);
  for (int I = 0; I < ns; ++I)
    for (int J = 0; J < ns; ++J) {
      int e = 0, f = 0;
      for (int k = 0; k < nc; ++k) bra[e++] = ket[f++] = k;
      for (int p = 0; p < S.nact; ++p) {
        if (str[I] >> p & 1) bra[e++] = nc + p;
        if (str[J] >> p & 1) ket[f++] = nc + p;
      }
      for (int r = 0; r < nel; ++r)
        for (int q = 0; q < nel; ++q)
          a[(size_t)r * nel + q] = (1.0 - x) * ov[bra[r] + (size_t)nf * ket[q]] + (bra[r] == ket[q] ? x : 0.0);
      out[(size_t)I * ns + J] = nel ? host_det(nel, a) : 1.0;
    }
}

}  // namespace
}  // extern "C++"

// P[k], k = 0..N: probability that exactly k electrons are outside rad
static int ipx_distribution(td1c_ctx* c, double rad, const double* lfield, const double* d_wfn, const double* d_cic,
                     std::vector<double>& P, hipStream_t s) {
  Setup& S = c->S;
  const int N = S.nel[2];
  if (S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (S.nact > 0 && !d_cic) return fail(-1, "null CI vector");
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  enqueue_orbin(c, c->lf_dev.p, reinterpret_cast<const double2*>(d_wfn), s);
  int cut = S.nr;
  for (int i = 0; i < S.nr; ++i)
    if (S.xr[i] > rad) {
      cut = i + 1;
      break;
    }
  const int nf = S.nfun;
  launch_window_ovlp(c->orb, c->mval.p, nf, S.nr, S.L1, cut, c->ovlp.p, s);
  HIP_TRY(hipGetLastError());
  std::vector<std::complex<double>> ov((size_t)nf * nf), cic(S.nact > 0 ? S.lcic : 1, 1.0);
  HIP_TRY(hipMemcpyAsync(ov.data(), c->ovlp.p, ov.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  if (S.nact > 0)
    HIP_TRY(hipMemcpyAsync(cic.data(), d_cic, cic.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // strings and the full alpha x beta coefficient table C[ib][ia]
  const std::vector<uint32_t> one{0u};
  const std::vector<uint32_t>& sa = S.nact > 0 ? S.str_a : one;
  const std::vector<uint32_t>& sb = S.nact > 0 ? S.str_b : one;
  const int nsa = (int)sa.size(), nsb = (int)sb.size();
  std::vector<std::complex<double>> C((size_t)nsa * nsb, 0.0);
  for (int ib = 0; ib < nsb; ++ib)
    for (int ia = 0; ia < nsa; ++ia) {
      const int d = S.nact > 0 ? S.det_of[(size_t)ib * nsa + ia] : 0;
      if (d >= 0) C[(size_t)ib * nsa + ia] = cic[d];
    }
  const double two_pi = 6.283185307179586476925286766559;
  std::vector<std::complex<double>> F(N + 1), da, db, X((size_t)nsb * nsa);
  for (int q = 0; q <= N; ++q) {
    const std::complex<double> x = std::polar(1.0, two_pi * q / (N + 1));
    spin_dets(S, ov, sa, S.nel[0], x, da);
    spin_dets(S, ov, sb, S.nel[1], x, db);
    // X[ib][ja] = sum_ia conj(C[ib][ia]) da[ia][ja]
    for (int ib = 0; ib < nsb; ++ib)
      for (int ja = 0; ja < nsa; ++ja) {
        std::complex<double> t = 0.0;
        for (int ia = 0; ia < nsa; ++ia) t += std::conj(C[(size_t)ib * nsa + ia]) * da[(size_t)ia * nsa + ja];
        X[(size_t)ib * nsa + ja] = t;
      }
    // F = sum_{ib,ja} X[ib][ja] sum_jb db[ib][jb] C[jb][ja]
    std::complex<double> f = 0.0;
    for (int ib = 0; ib < nsb; ++ib)
      for (int ja = 0; ja < nsa; ++ja) {
        std::complex<double> t = 0.0;
        for (int jb = 0; jb < nsb; ++jb) t += db[(size_t)ib * nsb + jb] * C[(size_t)jb * nsa + ja];
        f += X[(size_t)ib * nsa + ja] * t;
      }
    F[q] = f;
  }
  P.assign(N + 1, 0.0);
  for (int k = 0; k <= N; ++k) {
    std::complex<double> t = 0.0;
    for (int q = 0; q <= N; ++q) t += F[q] * std::polar(1.0, -two_pi * (double)q * k / (N + 1));
    P[k] = t.real() / (N + 1);
  }
  return 0;
}

int td1c_ipx(td1c_ctx* c, int max_ipx, double rad_ipx, const double* lfield, const double* d_wfn, const double* d_cic,
             double* ipx, void* stream) {
  if (!c || !lfield || !d_wfn || !ipx) return fail(-1, "null argument");
  if (max_ipx < 0 || max_ipx > 8 || max_ipx > c->S.nel[2])
    return fail(-1, "max_ipx must be in [0, min(8, electrons)]");
  std::vector<double> P;
  const int rc = ipx_distribution(c, rad_ipx, lfield, d_wfn, d_cic, P, (hipStream_t)stream);
  if (rc) return rc;
  for (int k = 0; k <= max_ipx; ++k) ipx[k] = P[k];
  return 0;
}

// hprod_ipd / ormas_ipd (src/hprod/hprod_ipx.F90:21-39; src/ormas/ormas_ipd.F90:2-123):
// S1 = sum_k k P(k) and S2 = sum_k k(k-1)/2 P(k), the first two factorial
// moments of the count outside rad_ipx (the reference forms them from the
// 1- and 2-RDM), give p0 = 1 - S1 + S2, p1 = S1 - 2 S2, p2 = S2.
int td1c_ipd(td1c_ctx* c, int max_ipd, double rad_ipx, const double* lfield, const double* d_wfn, const double* d_cic,
             double* ipd, void* stream) {
  if (!c || !lfield || !d_wfn || !ipd) return fail(-1, "null argument");
  if (max_ipd < 0 || max_ipd > 2 || max_ipd > c->S.nel[2])
    return fail(-1, "max_ipd must be in [0, min(2, electrons)]");
  std::vector<double> P;
  const int rc = ipx_distribution(c, rad_ipx, lfield, d_wfn, d_cic, P, (hipStream_t)stream);
  if (rc) return rc;
  double s1 = 0.0, s2 = 0.0;
  for (size_t k = 0; k < P.size(); ++k) {
    const double kk = (double)k;
    s1 += kk * P[k];
    s2 += 0.5 * kk * (kk - 1.0) * P[k];
  }
  const double p[3] = {1.0 - s1 + s2, s1 - 2.0 * s2, s2};
  for (int k = 0; k <= max_ipd; ++k) ipd[k] = p[k];
  return 0;
}

int td1c_mask(td1c_ctx* c, double* d_wfn, void* stream) {
  if (!c || !d_wfn) return fail(-1, "null argument");
  launch_mask(reinterpret_cast<double2*>(d_wfn), c->mask.p, c->S.nr, c->S.L1, c->S.nfcore, c->S.nfun,
              (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_orth(td1c_ctx* c, double* d_wfn, void* stream) {
  if (!c || !d_wfn) return fail(-1, "null argument");
  Setup& S = c->S;
  double2* w = reinterpret_cast<double2*>(d_wfn);
  // wfn_orth_cic normalises the CI vector whatever the active space (one
  // coefficient without active orbitals)
  launch_orth(w, w + (size_t)c->nbas * S.nfun, c->mval.p, S.nr, S.L1, S.nfcore, S.nfun, S.lcic, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return 0;
}

// clwfn::write (src/wfn/clwfn.cpp:268-302): the .orb file holds nrad, lmax1,
// mmax1, nfun ("%10d" lines) then every orbital coefficient as "%25.15le%25.15le"
// in [orbital][l][r] order; the .cic file holds lcic then the CI coefficients.
int td1c_wfn_write(const char* orb_path, const char* cic_path, int nrad, int lmax1, int mmax1, int nfun, int lcic,
                   const double* wfn) {
  if (!orb_path || !cic_path || !wfn) return fail(-1, "null argument");
  if (nrad < 2 || lmax1 < 0 || mmax1 < 0 || nfun < 0 || lcic < 0) return fail(-1, "bad sizes");
  FILE* fo = std::fopen(orb_path, "w");
  if (!fo) return fail(-5, std::string("cannot open ") + orb_path);
  std::fprintf(fo, "%10d\n%10d\n%10d\n%10d\n", nrad, lmax1, mmax1, nfun);
  const size_t no = (size_t)nfun * (lmax1 + 1) * (nrad - 1);
  for (size_t q = 0; q < no; ++q) std::fprintf(fo, "%25.15le%25.15le\n", wfn[2 * q], wfn[2 * q + 1]);
  const bool ok_o = std::fclose(fo) == 0;
  FILE* fc = std::fopen(cic_path, "w");
  if (!fc) return fail(-5, std::string("cannot open ") + cic_path);
  std::fprintf(fc, "%10d\n", lcic);
  for (int d = 0; d < lcic; ++d) std::fprintf(fc, "%25.15le%25.15le\n", wfn[2 * (no + d)], wfn[2 * (no + d) + 1]);
  const bool ok_c = std::fclose(fc) == 0;
  if (!ok_o || !ok_c) return fail(-5, "write error");
  return 0;
}

// clwfn::read (src/wfn/clwfn.cpp:156-205): the wave function is cleared, the
// file's orbitals are read and those inside (nfun, lmax1, nrad) kept (a
// smaller file leaves zeros, a larger one is truncated); the CI vector is read
// when the file's lcic equals lcic, else determinant cic0 is set to 1
// (ormas_cic0, src/ormas/ormas_cic0.F90:24; left zero when cic0 < 0, i.e.
// without active orbitals) and 1 is returned.
int td1c_wfn_read(const char* orb_path, const char* cic_path, int nrad, int lmax1, int nfun, int lcic, int cic0,
                  double* wfn) {
  if (!orb_path || !cic_path || !wfn) return fail(-1, "null argument");
  if (nrad < 2 || lmax1 < 0 || nfun < 0 || lcic < 0) return fail(-1, "bad sizes");
  if (cic0 >= lcic) return fail(-1, "cic0 outside the CI vector");
  const size_t no = (size_t)nfun * (lmax1 + 1) * (nrad - 1);
  std::fill(wfn, wfn + 2 * (no + lcic), 0.0);
  FILE* fo = std::fopen(orb_path, "r");
  if (!fo) return fail(-5, std::string("cannot open ") + orb_path);
  int nrad0, lmax0, mmax0, nfun0;
  if (std::fscanf(fo, "%d %d %d %d", &nrad0, &lmax0, &mmax0, &nfun0) != 4) {
    std::fclose(fo);
    return fail(-5, std::string("bad header in ") + orb_path);
  }
  for (int f = 0; f < nfun0; ++f)
    for (int l = 0; l <= lmax0; ++l)
      for (int r = 1; r < nrad0; ++r) {
        double re, im;
        if (std::fscanf(fo, "%le %le", &re, &im) != 2) {
          std::fclose(fo);
          return fail(-5, std::string("short file ") + orb_path);
        }
        if (f < nfun && l <= lmax1 && r < nrad) {
          const size_t q = ((size_t)f * (lmax1 + 1) + l) * (nrad - 1) + (r - 1);
          wfn[2 * q] = re;
          wfn[2 * q + 1] = im;
        }
      }
  std::fclose(fo);
  FILE* fc = std::fopen(cic_path, "r");
  if (!fc) return fail(-5, std::string("cannot open ") + cic_path);
  int lcic0;
  if (std::fscanf(fc, "%d", &lcic0) != 1) {
    std::fclose(fc);
    return fail(-5, std::string("bad header in ") + cic_path);
  }
  int rc = 0;
  if (lcic0 != lcic) {
    if (cic0 >= 0) wfn[2 * (no + cic0)] = 1.0;
    rc = 1;
  } else {
    for (int d = 0; d < lcic; ++d)
      if (std::fscanf(fc, "%le %le", &wfn[2 * (no + d)], &wfn[2 * (no + d) + 1]) != 2) {
        std::fclose(fc);
        return fail(-5, std::string("short file ") + cic_path);
      }
  }
  std::fclose(fc);
  return rc;
}

// clbas::proj (src/bas/clbas.cpp:280-332) as guess_type = "read_proj" uses it
// (src/main/guess.cpp:37-47): every interior point r of the target grid gets
// sqrt(w_target(r)) * sum_m c_m L_m(r), L_m the Lagrange polynomials of the
// source element holding r (clfedvr::get_val, src/fedvr/clfedvr.cpp:651-675,
// cldvr::get_val, src/dvr/cldvr.cpp:223-235) over its stored points, its right
// bridge included; the source coefficients c_m enter as they are (without
// 1/sqrt(w_source), as in the reference).  The source element is the first
// with r + 1e-10 < its right edge (clfedvr::get_ife, :628-638); points beyond
// the source grid stay zero.  The CI vector is copied (zcopy of the target's
// length, guess.cpp:47); orthonormalisation (clwfn::orth) is td1c_orth's.
int td1c_wfn_proj(const td1c_problem* from, const double* wfn_from, const td1c_problem* to, double* wfn_to) {
  if (!from || !to || !wfn_from || !wfn_to) return fail(-1, "null argument");
  for (const td1c_problem* p : {from, to}) {
    if (p->nfe < 2 || p->ndvr < 2 || !p->edges) return fail(-1, "need nfe >= 2, ndvr >= 2 and element edges");
    if (p->lmax1 < 0) return fail(-1, "bad lmax1");
    for (int e = 0; e < p->nfe; ++e)
      if (!(p->edges[e + 1] > p->edges[e])) return fail(-1, "element edges must increase");
  }
  Setup A, B;
  if (int rc = ci_setup(from, A)) return rc;
  if (int rc = ci_setup(to, B)) return rc;
  const int n1 = from->ndvr, nfe1 = from->nfe, np1 = n1 + 1;
  const int n2 = to->ndvr, nfe2 = to->nfe;
  const int nr1 = n1 * nfe1 - 1, nr2 = n2 * nfe2 - 1;
  const int L1a = from->lmax1 + 1, L1b = to->lmax1 + 1;
  std::vector<double> ex1, ew1, xr1, wr1, ex2, ew2, xr2, wr2;
  fedvr_nodes(nfe1, n1, from->edges, ex1, ew1, xr1, wr1);
  fedvr_nodes(nfe2, n2, to->edges, ex2, ew2, xr2, wr2);
  const size_t no1 = (size_t)A.nfun * L1a * nr1, no2 = (size_t)B.nfun * L1b * nr2;
  std::fill(wfn_to, wfn_to + 2 * (no2 + B.lcic), 0.0);
  const int nf = std::min(A.nfun, B.nfun), nl = std::min(L1a, L1b);
  // one target point: the source points it reads and their weights
  std::vector<int> src(np1);
  std::vector<double> val(np1);
  for (int i2 = 1; i2 <= nr2; ++i2) {
    const double r = xr2[i2], sw = std::sqrt(wr2[i2]);
    int e1 = -1;
    for (int e = 0; e < nfe1; ++e)
      if (r + 1e-10 < from->edges[e + 1]) { e1 = e; break; }
    if (e1 < 0) continue;
    const double* xe = &ex1[(size_t)e1 * np1];
    auto lagrange = [&](int m) {
      double v = 1.0;
      for (int k = 0; k <= n1; ++k)
        if (k != m) v *= (r - xe[k]) / (xe[m] - xe[k]);
      return v;
    };
    int nt = 0;
    for (int m = e1 == 0 ? 1 : 0; m < n1; ++m) { src[nt] = e1 * n1 + m; val[nt++] = lagrange(m) * sw; }
    if (e1 < nfe1 - 1) { src[nt] = (e1 + 1) * n1; val[nt++] = lagrange(n1) * sw; }
    for (int f = 0; f < nf; ++f)
      for (int l = 0; l < nl; ++l) {
        const double* c1 = wfn_from + 2 * ((size_t)f * L1a + l) * nr1;  // c1[2 (i1 - 1)]: point i1
        double* c2 = wfn_to + 2 * (((size_t)f * L1b + l) * nr2 + i2 - 1);
        for (int t = 0; t < nt; ++t) {
          c2[0] += c1[2 * (src[t] - 1)] * val[t];
          c2[1] += c1[2 * (src[t] - 1) + 1] * val[t];
        }
      }
  }
  const int nci = std::min(A.lcic, B.lcic);
  std::copy(wfn_from + 2 * no1, wfn_from + 2 * (no1 + nci), wfn_to + 2 * no2);
  return 0;
}

// clcrk tableaux for rk_order 1..4 (src/prop/clcrk.cpp:183-235)
static int rk_tableau(int order, int& ns, double* nodes, double* w, double (*a)[kMaxStages]) {
  for (int i = 0; i < kMaxStages; ++i) {
    nodes[i] = w[i] = 0;
    for (int j = 0; j < kMaxStages; ++j) a[i][j] = 0;
  }
  switch (order) {
    case 1: ns = 1; w[0] = 1; return 0;
    case 2: ns = 2; nodes[1] = 0.5; w[1] = 1; a[1][0] = 0.5; return 0;
    case 3: ns = 3; nodes[1] = 0.5; nodes[2] = 1; w[0] = 1.0 / 6; w[1] = 2.0 / 3; w[2] = 1.0 / 6;
            a[1][0] = 0.5; a[2][0] = -1; a[2][1] = 2; return 0;
    case 4: ns = 4; nodes[1] = 0.5; nodes[2] = 0.5; nodes[3] = 1;
            w[0] = 1.0 / 6; w[1] = 1.0 / 3; w[2] = 1.0 / 3; w[3] = 1.0 / 6;
            a[1][0] = 0.5; a[2][1] = 0.5; a[3][2] = 1; return 0;
    default: return -1;
  }
}

// One Runge-Kutta step of the whole wave function on stream s, the stage
// field values read from d_lf (device, [stage][9]).
// join: make s wait for the island stage B at the end (always, unless the
// caller enqueues further steps and joins after the last one: the CI data
// of consecutive steps stays on side stream 0, the orbitals on s)
// split_type = 1: the Crank-Nicolson factors of clh1itr::gen for a step of
// dtime (half steps of dtime / 2, h1itr_gen_cninv's factor (i) dtime / 4),
// built on the host and uploaded when dtime or the frozen core changed.
static int cn_prepare(td1c_ctx* c, double dtime) {
  Setup& S = c->S;
  if (S.isplit != 1) return fail(-1, "the one-body half steps need split_type = 1");
  if (c->cn_dt == dtime && c->cnG.p) return 0;
  if (S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (!(dtime != 0.0) || !std::isfinite(dtime)) return fail(-1, "the split step needs a finite nonzero dtime");
  CnTables T;
  std::string err;
  const double q = 0.25 * dtime;
  if (!build_cn(S, c->v2jfc_h.empty() ? nullptr : c->v2jfc_h.data(), S.icomp == 1 ? 0.0 : q, S.icomp == 1 ? q : 0.0,
                T, err))
    return fail(-5, err);
  if (c->cnG.p) HIP_TRY(hipDeviceSynchronize());  // earlier half steps may still read the old factors
  if (c->gexec) {  // a captured step holds the old factors' addresses
    HIP_TRY(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  auto up = [](DevBuf<double2>& b, const std::vector<double>& v) -> hipError_t {
    b.release();
    hipError_t e = b.alloc(std::max<size_t>(v.size() / 2, 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
  };
  HIP_TRY(up(c->cnG, T.G));
  c->cnG_h = std::move(T.G);  // host copy: batches share equal tables
  HIP_TRY(up(c->cnWl, T.Wl));
  HIP_TRY(up(c->cnWr, T.Wr));
  HIP_TRY(up(c->cnPa, T.pa));
  HIP_TRY(up(c->cnPg, T.pg));
  HIP_TRY(up(c->cnBinv, T.binv));
  c->cn_K = T.K;
  const int nd = S.nfun - S.nfcore;
  const size_t nrow = (size_t)std::max(nd, 1) * S.L1 * S.nr;
  if (!c->cn_x.p) HIP_TRY(c->cn_x.alloc(nrow));
  if (!c->cn_y.p) HIP_TRY(c->cn_y.alloc(nrow));
  if (!c->cn_flag.p) {
    HIP_TRY(c->cn_flag.alloc(2));
    HIP_TRY(hipMemset(c->cn_flag.p, 0, 2 * sizeof(int)));
  }
  if (!c->lf_cn.p) HIP_TRY(c->lf_cn.alloc(9));
  c->cn_npart1 = std::max(nd, 1) * S.L1;
  CnArgs probe{};
  probe.nfe = S.nfe; probe.ni = S.n - 1; probe.L1 = S.L1; probe.nfun = S.nfun; probe.nfcore = S.nfcore;
  const size_t npart = (size_t)c->cn_npart1 + cn_interior_blocks(probe);
  if (c->cn_part.n < npart) {
    c->cn_part.release();
    HIP_TRY(c->cn_part.alloc(npart));
  }
  c->cn_dt = dtime;
  return 0;
}

static CnArgs cn_args(td1c_ctx* c, const double* d_lf, double2* wfn) {
  Setup& S = c->S;
  CnArgs a{};
  a.nr = S.nr; a.n = S.n; a.nfe = S.nfe; a.ni = S.n - 1; a.nb = S.nfe - 1; a.K = c->cn_K; a.L1 = S.L1;
  a.nfun = S.nfun; a.nfcore = S.nfcore; a.mmax1 = S.mmax1; a.igauge = S.igauge; a.isplit = S.isplit;
  a.mval = c->mval.p;
  a.band_k = c->band_k.p; a.band_k0 = c->band_k0.p; a.band_p = c->band_p.p; a.tdiag = c->tdiag.p;
  a.xr = c->xr.p; a.alph = c->alph.p; a.lf = d_lf;
  a.G = c->cnG.p; a.gstride = S.nfe; a.Wl = c->cnWl.p; a.Wr = c->cnWr.p; a.pa = c->cnPa.p; a.pg = c->cnPg.p; a.binv = c->cnBinv.p;
  a.wfn = wfn; a.x = c->cn_x.p; a.y = c->cn_y.p;
  const double q = 0.25 * c->cn_dt;
  a.zfac1 = S.icomp == 1 ? make_double2(0.0, -q) : make_double2(-q, 0.0);
  a.thresh = c->cn_thresh;
  a.flag = c->cn_flag.p;
  a.part = c->cn_part.p;
  a.npart1 = c->cn_npart1;
  return a;
}

// maxcyc more terms (mode 1) of a series whose first term is enqueued.
// A captured or recorded step holds every term (the converged ones return
// at once).  Under stream launches, after the second term and every two
// terms after that, the host reads the flag back (the call waits for the
// device to get there) and stops enqueuing once the series has converged:
// the later terms would all return at once, and each would still cost its
// launches.
static int enqueue_series(td1c_ctx* c, const CnArgs& a, const OneBodyElArgs* obp, int maxcyc, hipStream_t s) {
  if (c->capturing || active_recorder()) {
    for (int k = 0; k < maxcyc; ++k) launch_cn_stage(a, obp, c->ob_shape, c->ob_lc, 1, s);
    return 0;
  }
  if (!c->cn_flag_host) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c->cn_flag_host), sizeof(int)));
  int k = 0;
  while (k < maxcyc) {
    const int chunk = k == 0 ? 1 : 2;
    for (int j = 0; j < chunk && k < maxcyc; ++j, ++k) launch_cn_stage(a, obp, c->ob_shape, c->ob_lc, 1, s);
    if (k >= maxcyc) break;
    launch_cn_check(a, s);
    HIP_TRY(hipMemcpyAsync(c->cn_flag_host, a.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (*c->cn_flag_host) break;
  }
  return 0;
}

// One clh1itr half step (h1itr_prop) on the orbitals of wfn with the device
// field d_lf: the first term, then cn_maxcyc more terms that return at once
// after convergence.
static int enqueue_h1itr(td1c_ctx* c, const double* d_lf, double2* wfn, hipStream_t s) {
  const CnArgs a = cn_args(c, d_lf, wfn);
  // the right-hand sides by the element one-body kernel (ndvr = 20): over
  // the dynamical orbitals, with the field of this half step
  OneBodyElArgs ob = onebody_el_args(c, d_lf);
  ob.nfun = c->S.nfun - c->S.nfcore;
  ob.nfcore = 0;
  ob.fshift = c->S.nfcore;
  ob.isplit = c->S.isplit;
  ob.cv = a.zfac1;
  ob.stop = a.flag;
  const OneBodyElArgs* obp = c->S.n == 20 ? &ob : nullptr;
  TIMED("h1itr");
  launch_cn_stage(a, obp, c->ob_shape, c->ob_lc, 0, s);
  return enqueue_series(c, a, obp, c->cn_maxcyc, s);
}

static int enqueue_rk_step(td1c_ctx* c, int rk_order, double dtime, const double* d_lf, double2* wfn, int apply_mask,
                           hipStream_t s, bool join = true) {
  int ns;
  double nodes[kMaxStages], w[kMaxStages], a[kMaxStages][kMaxStages];
  if (rk_tableau(rk_order, ns, nodes, w, a)) return fail(-1, "rk_order must be 1..4");
  Setup& S = c->S;
  const size_t so = (size_t)c->nbas * S.nfun, sz = so + S.lcic;
  double2* w0 = c->rk_w0.p;
  double2* t = c->rk_t.p;
  // stage 0 reads the state in place; its dtorb and island kernels read the
  // step-start state before updating it and leave a copy in w0 for the
  // later stages
  const double2* tin = wfn;
  for (int is = 0; is < ns; ++is) {
    double2* k = c->rk_k.p + sz * is;
    CiStage cs;
    cs.kout = k + so;
    cs.acc = wfn + so;
    cs.dtw = dtime * w[is];
    if (is + 1 < ns) {
      cs.next = t + so;
      cs.w0 = is == 0 ? wfn + so : w0 + so;
      if (is == 0) cs.w0out = w0 + so;
      cs.nprev = is;
      for (int js = 0; js < is; ++js) {
        cs.kprev[js] = c->rk_k.p + sz * js + so;
        cs.cprev[js] = dtime * a[is + 1][js];
      }
      cs.cnew = dtime * a[is + 1][is];
    }
    int rc = enqueue_htotx(c, 1.0, d_lf + 9 * is, tin, tin + so, cs, s);
    if (rc) return rc;
    DtorbArgs d = dtorb_args(c, 1.0);
    d.kout = k;
    d.acc = wfn;
    d.dtw = dtime * w[is];
    d.mask = (apply_mask && is == ns - 1) ? c->mask.p : nullptr;
    if (is + 1 < ns) {
      d.next = t;
      d.w0 = is == 0 ? wfn : w0;
      if (is == 0) d.w0out = w0;
      d.nprev = is;
      for (int js = 0; js < is; ++js) {
        d.kprev[js] = c->rk_k.p + sz * js;
        d.cprev[js] = dtime * a[is + 1][js];
      }
      d.cnew = dtime * a[is + 1][is];
    }
    { TIMED("dtorb_rk"); launch_dtorb(d, s); }
    if (is + 1 < ns) tin = t;
  }
  if (join) HIP_TRY(join_island(c, s));
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_rk_prop(td1c_ctx* c, int rk_order, double dtime, const double* lfield_stages, double* d_wfn, int apply_mask,
                 void* stream) {
  if (!c || !lfield_stages || !d_wfn) return fail(-1, "null argument");
  if (c->S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  int ns;
  double nodes[kMaxStages], w[kMaxStages], a[kMaxStages][kMaxStages];
  if (rk_tableau(rk_order, ns, nodes, w, a)) return fail(-1, "rk_order must be 1..4");
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield_stages, 9 * ns * sizeof(double), hipMemcpyHostToDevice, s));
  return enqueue_rk_step(c, rk_order, dtime, c->lf_dev.p, reinterpret_cast<double2*>(d_wfn), apply_mask, s);
}

// init_split2 with split_type 0 (src/main/init.cpp:178-240): imaginary-time
// relaxation.  Each cycle evaluates the energy at lfield (hprod_energy), stops
// when it changed by less than tolene, else takes one rk_order step of dstep
// (clcrk::prop, no mask) and orthonormalises (clwfn::orth, schmidt).  *ene
// receives the last energy, *ncyc the last cycle index.
int td1c_h1itr_gen(td1c_ctx* c, double dtime, int maxcyc, double thresh) {
  if (!c) return fail(-1, "null argument");
  if (c->S.isplit != 1) return fail(-1, "td1c_h1itr_gen needs a context with split_type = 1");
  if (maxcyc < 0 || !(thresh >= 0.0)) return fail(-1, "h1itr_maxcyc must be >= 0 and h1itr_thresh >= 0");
  c->cn_maxcyc = maxcyc;
  c->cn_thresh = thresh;
  if (c->gexec) {  // a captured step holds the old cycle count
    HIP_TRY(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  c->cn_dt = 0.0;
  return cn_prepare(c, dtime);
}

int td1c_h1itr_prop(td1c_ctx* c, const double* lfield, double* d_wfn, void* stream) {
  if (!c || !lfield || !d_wfn) return fail(-1, "null argument");
  if (c->S.isplit != 1) return fail(-1, "td1c_h1itr_prop needs a context with split_type = 1");
  if (!c->cnG.p || c->cn_dt == 0.0) return fail(-4, "td1c_h1itr_gen first");
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(c->lf_cn.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  int rc = enqueue_h1itr(c, c->lf_cn.p, reinterpret_cast<double2*>(d_wfn), s);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_init_split2(td1c_ctx* c, int rk_order, double dstep, double tolene, int maxcyc, const double* lfield,
                     double* d_wfn, double* ene, int* ncyc, void* stream) {
  if (!c || !lfield || !d_wfn || !ene || !ncyc) return fail(-1, "null argument");
  if (c->S.icomp != 0) return fail(-1, "imaginary-time relaxation needs a context with icomp = 0");
  if (c->S.isplit != 0) return fail(-1, "td1c_init_split2 runs split_type = 0 only");
  if (maxcyc < 0 || !(dstep > 0.0)) return fail(-1, "maxcyc must be >= 0 and dstep > 0");
  int ns;
  double nodes[kMaxStages], w[kMaxStages], a[kMaxStages][kMaxStages];
  if (rk_tableau(rk_order, ns, nodes, w, a)) return fail(-1, "rk_order must be 1..4");
  std::vector<double> lfs((size_t)9 * ns);
  for (int st = 0; st < ns; ++st) std::copy(lfield, lfield + 9, lfs.begin() + 9 * st);
  const size_t no = (size_t)c->nbas * c->S.nfun;
  double* scratch = reinterpret_cast<double*>(c->rk_t.p);
  double e1 = -1.0e10, e[5] = {0, 0, 0, 0, 0};
  int last = 0, rc = 0;
  for (int icyc = 0; icyc < maxcyc; ++icyc) {
    rc = td1c_htot(c, dstep, lfield, d_wfn, d_wfn + 2 * no, scratch, scratch + 2 * no, stream);
    if (rc == 0) rc = td1c_get_energies(c, e);
    if (rc) return rc;
    last = icyc;
    if (std::fabs(e[4] - e1) < tolene) break;
    e1 = e[4];
    rc = td1c_rk_prop(c, rk_order, dstep, lfs.data(), d_wfn, 0, stream);
    if (rc == 0) rc = td1c_orth(c, d_wfn, stream);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  *ene = e[4];
  *ncyc = last;
  return 0;
}

// hprod_chkconv (src/hprod/hprod_chkconv.F90:2-63): out[0] the energy of an
// hprod, out[1] |hcic| / lcic, and from the generalised Fock vectors g_i of
// the dynamical orbitals (hprod_gfock: (h0 + h1 + g + v2) phi_i, here the
// hprod's own h01, g and v2 of that orbital) projected on the complement of
// all orbitals (hprod_projall, inner false: g_i -= sum_j phi_j <phi_j|g_i>)
// and their overlaps over the whole box (hprod_mkovlp): out[2] =
// sqrt(sum_i <g_i|g_i>) / nfun, out[3] = sqrt(2 sum_dcore <g_i|g_i> +
// sum_act <g_a|g_b> den1(b, a)).  Synchronises the stream.
int td1c_chkconv(td1c_ctx* c, const double* lfield, const double* d_wfn, double out[4], void* stream) {
  if (!c || !lfield || !d_wfn || !out) return fail(-1, "null argument");
  Setup& S = c->S;
  if (S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (S.nradfc != S.nr) return fail(-1, "td1c_chkconv needs nradfc = nrad-1");
  hipStream_t s = (hipStream_t)stream;
  const int nf = S.nfun, nd = nf - S.nfcore, NF = nf + nd;
  const size_t nb = c->nbas, so = nb * nf, ndn = nb * nd;
  if (!c->chk_buf.p) {
    HIP_TRY(c->chk_buf.alloc((size_t)NF * nb + std::max(S.lcic, 1)));
    HIP_TRY(c->chk_ov.alloc((size_t)NF * NF));
    std::vector<int> mv(NF);
    for (int i = 0; i < nf; ++i) mv[i] = S.mval[i];
    for (int i = 0; i < nd; ++i) mv[nf + i] = S.mval[S.nfcore + i];
    HIP_TRY(c->chk_mval.upload(mv));
  }
  const double2* wfn = reinterpret_cast<const double2*>(d_wfn);
  double2* cb = c->chk_buf.p;
  double2* hcic = cb + (size_t)NF * nb;
  HIP_TRY(hipMemcpyAsync(c->lf_dev.p, lfield, 9 * sizeof(double), hipMemcpyHostToDevice, s));
  CiStage cs;
  cs.kout = hcic;
  if (int rc = enqueue_htotx(c, 1.0, c->lf_dev.p, wfn, wfn + so, cs, s)) return rc;
  HIP_TRY(join_island(c, s));
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  double2* g = cb + so;
  launch_zaxpby(so, one, c->orb, zero, nullptr, cb, s);
  launch_zaxpby(ndn, one, c->v2orb, one, c->gorb, g, s);
  launch_zaxpby(ndn, one, g, one, c->h01 + (size_t)S.nfcore * nb, g, s);
  launch_window_ovlp(cb, c->chk_mval.p, NF, S.nr, S.L1, S.nr, c->chk_ov.p, s);
  std::vector<double2> ov((size_t)NF * NF);
  HIP_TRY(hipMemcpyAsync(ov.data(), c->chk_ov.p, ov.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int i = 0; i < nd; ++i)  // g_i -= sum_j phi_j fmat(j, i), fmat(j, i) = <phi_j|g_i>
    for (int j = 0; j < nf; ++j) {
      if (S.mval[j] != S.mval[S.nfcore + i]) continue;
      const double2 f = ov[j + (size_t)NF * (nf + i)];
      launch_zaxpby(nb, make_double2(-f.x, -f.y), cb + (size_t)j * nb, one, g + (size_t)i * nb, g + (size_t)i * nb, s);
    }
  launch_window_ovlp(g, c->chk_mval.p + nf, nd, S.nr, S.L1, S.nr, c->chk_ov.p, s);
  std::vector<double2> go((size_t)nd * nd), hc(std::max(S.lcic, 1));
  std::vector<double2> den1((size_t)std::max(S.nact * S.nact, 1));
  HIP_TRY(hipMemcpyAsync(go.data(), c->chk_ov.p, go.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(hc.data(), hcic, hc.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  if (S.nact > 0)
    HIP_TRY(hipMemcpyAsync(den1.data(), c->den1.p, den1.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double e[5];
  if (int rc = td1c_get_energies(c, e)) return rc;
  out[0] = e[4];
  double hn = 0.0;
  for (int k = 0; k < S.lcic; ++k) hn += hc[k].x * hc[k].x + hc[k].y * hc[k].y;
  out[1] = std::sqrt(hn) / S.lcic;
  double c2 = 0.0, c3 = 0.0;
  for (int i = 0; i < nd; ++i) c2 += go[i + (size_t)nd * i].x;
  for (int i = 0; i < S.ndcore; ++i) c3 += 2.0 * go[i + (size_t)nd * i].x;
  const int a0 = S.ndcore;  // the first active orbital among the dynamical ones
  for (int ia = 0; ia < S.nact; ++ia)
    for (int ja = 0; ja < S.nact; ++ja) {
      const double2 o = go[(a0 + ia) + (size_t)nd * (a0 + ja)], d = den1[ja + (size_t)S.nact * ia];
      c3 += o.x * d.x - o.y * d.y;
    }
  out[2] = std::sqrt(c2) / nf;
  out[3] = std::sqrt(c3);
  return 0;
}

// The stage nodes of one propagation step: the RK nodes, and with
// split_type = 1 the two half-step field times t + dt/4 and t + 3dt/4 as two
// more nodes (ns = the RK stages).
static int step_nodes(td1c_ctx* c, int rk_order, StageNodes& sn, int& ns) {
  double nodes[kMaxStages], w[kMaxStages], a[kMaxStages][kMaxStages];
  if (rk_tableau(rk_order, ns, nodes, w, a)) return fail(-1, "rk_order must be 1..4");
  sn = StageNodes{};
  sn.ns = ns;
  for (int i = 0; i < ns; ++i) sn.c[i] = nodes[i];
  if (c->S.isplit == 1) {
    if (ns + 2 > kMaxStages) return fail(-1, "too many stages for the split step");
    sn.c[ns] = 0.25;
    sn.c[ns + 1] = 0.75;
    sn.ns = ns + 2;
  }
  return 0;
}

// One step of td1c_propagate after its stage-field launch.  split_type = 1:
// tdse_split2's step is clh1itr::prop(STEP1) with the field at t + dt/4, the
// RK step of the split hprod, clh1itr::prop(STEP2) at t + 3dt/4, then the
// mask (src/main/tdse.cpp:225-266, src/prop/clh1itr.cpp:61-91).
static int enqueue_prop_step(td1c_ctx* c, int rk_order, int ns, double dtime, double2* wfn, int apply_mask,
                             hipStream_t st, bool join) {
  const bool split = c->S.isplit == 1;
  if (split) {
    int rc = enqueue_h1itr(c, c->lf_dev.p + 9 * ns, wfn, st);
    if (rc) return rc;
  }
  int rc = enqueue_rk_step(c, rk_order, dtime, c->lf_dev.p, wfn, split ? 0 : apply_mask, st, join);
  if (rc) return rc;
  if (split) {
    rc = enqueue_h1itr(c, c->lf_dev.p + 9 * (ns + 1), wfn, st);
    if (rc) return rc;
    if (apply_mask) launch_mask(wfn, c->mask.p, c->S.nr, c->S.L1, c->S.nfcore, c->S.nfun, st);
  }
  return 0;
}

static int check_field(const td1c_field* field) {
  if (field->env_type < 0 || field->env_type > 2 || field->gauge < 0 || field->gauge > 1)
    return fail(-1, "bad field");
  return 0;
}

int td1c_propagate(td1c_ctx* c, const td1c_field* field, int rk_order, double t0, double dtime, int nstep,
                   double* d_wfn, int apply_mask, void* stream) {
  if (!c || !field || !d_wfn) return fail(-1, "null argument");
  if (c->S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (check_field(field)) return -1;
  if (nstep < 0) return fail(-1, "nstep must be >= 0");
  int ns;
  StageNodes sn;
  if (int rc = step_nodes(c, rk_order, sn, ns)) return rc;
  if (nstep == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const Pulse p = make_pulse(field);
  if (c->S.isplit == 1) {
    int rc = cn_prepare(c, dtime);
    if (rc) return rc;
  }
  double2* wfn = reinterpret_cast<double2*>(d_wfn);
  if (!c->use_graph) {
    // stream launches: the hprod's side-stream branches then run concurrently
    launch_clock_set(c->clock.p, t0, dtime, s);
    for (int i = 0; i < nstep; ++i) {
      launch_field_stages(p, c->clock.p, sn, c->lf_dev.p, s);
      int rc = enqueue_prop_step(c, rk_order, ns, dtime, wfn, apply_mask, s, i == nstep - 1);
      if (rc) return rc;
    }
    HIP_TRY(hipGetLastError());
    return 0;
  }
  const bool same = c->gexec && c->g_order == rk_order && c->g_mask == apply_mask && c->g_wfn == d_wfn &&
                    c->g_dt == dtime && std::memcmp(&c->g_pulse, &p, sizeof(Pulse)) == 0;
  if (!same) {
    if (c->gexec) {
      HIP_TRY(hipGraphExecDestroy(c->gexec));
      c->gexec = nullptr;
    }
    if (!c->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    launch_field_stages(p, c->clock.p, sn, c->lf_dev.p, c->cap_stream);
    int rc = enqueue_prop_step(c, rk_order, ns, dtime, wfn, apply_mask, c->cap_stream, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->cap_stream, &g);
    c->capturing = false;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(-3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->gexec = nullptr;
      return fail(-3, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c->g_order = rk_order;
    c->g_mask = apply_mask;
    c->g_wfn = d_wfn;
    c->g_dt = dtime;
    c->g_pulse = p;
  }
  launch_clock_set(c->clock.p, t0, dtime, s);
  for (int i = 0; i < nstep; ++i) HIP_TRY(hipGraphLaunch(c->gexec, s));
  HIP_TRY(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// cletdrb: exponential Rosenbrock RK4 with Pade forms of the phi functions
// (cletdrb::prop4_orb, src/prop/cletdrb.cpp:321-434; clh1rat,
// src/prop/clh1rat.cpp; h1rat_gen / h1rat_prod, src/prop/h1rat_gen.F90,
// src/prop/h1rat_prod.F90), with the defaults of cletdrb::gen (:25-92):
// etd_rk_order 4, the [2/3] forms of phi1, phi2 and phi3, etd_time_of_h 0,
// etd_cisplit -1, split_type 1.  A step is four split hprods and seven
// rational functions of h1: numerators by two applications of h1, each
// denominator factor cd0 + cd1 (T + V) by the element Schur solve of
// 1 + (cd1/cd0) T with the Neumann series in V of the half steps.
// ---------------------------------------------------------------------------
namespace {

enum { kPhi1Dt1, kPhi1Dt2, kPhid1Dt1, kPhid1Dt2, kPhi2Dt1, kPhi2Dt2, kPhi3Dt1, kEtdRats };
enum { kExp1, kExp2, kGfun, kF0, kF1, kF2 };  // cletdrk4's functions

// The [2/3] forms: numerator coefficients, the sign of the first denominator
// factor and the denominator roots (gen_phi1 :636-771, gen_phid1 :773-886,
// gen_phi2 :888-1035, gen_phi3 :1037-1185).  fam 0 phi1, 1 phid1 (phi2
// over phi1's denominator), 2 phi2, 3 phi3.
void etd_form(int fam, double& dfac, double num[3], std::complex<double> root[3]) {
  static const double kRoots[3][3][2] = {
      {{4.644370709252173768, 0.0}, {3.677814645373918445, 3.508761919567442433},
       {3.677814645373918445, -3.508761919567442433}},
      {{5.648485971016884655, 0.0}, {4.675757014491553676, 3.913489560603714335},
       {4.675757014491553676, -3.913489560603714335}},
      {{6.651316808950019421, 0.0}, {5.674341595524996507, 4.279971984629763249},
       {5.674341595524996507, -4.279971984629763249}}};
  static const double kNum[4][3] = {{120.0, 0.0, 2.0},
                                    {60.0, -10.0, 1.0},
                                    {420.0 / 4.0, -40.0 / 4.0, 5.0 / 4.0},
                                    {1120.0 / 20.0, -140.0 / 20.0, 11.0 / 20.0}};
  const int rs = fam <= 1 ? 0 : fam - 1;
  dfac = -1.0;
  for (int k = 0; k < 3; ++k) {
    num[k] = kNum[fam][k];
    root[k] = std::complex<double>(kRoots[rs][k][0], kRoots[rs][k][1]);
  }
}

double2 d2(std::complex<double> z) { return make_double2(z.real(), z.imag()); }

// the Crank-Nicolson-form work buffers shared with the half steps
int cn_buffers(td1c_ctx* c) {
  Setup& S = c->S;
  const int nd = S.nfun - S.nfcore;
  const size_t nrow = (size_t)std::max(nd, 1) * S.L1 * S.nr;
  if (!c->cn_x.p) HIP_TRY(c->cn_x.alloc(nrow));
  if (!c->cn_y.p) HIP_TRY(c->cn_y.alloc(nrow));
  if (!c->cn_flag.p) {
    HIP_TRY(c->cn_flag.alloc(2));
    HIP_TRY(hipMemset(c->cn_flag.p, 0, 2 * sizeof(int)));
  }
  c->cn_npart1 = std::max(nd, 1) * S.L1;
  CnArgs probe{};
  probe.nfe = S.nfe; probe.ni = S.n - 1; probe.L1 = S.L1; probe.nfun = S.nfun; probe.nfcore = S.nfcore;
  const size_t npart = (size_t)c->cn_npart1 + cn_interior_blocks(probe);
  if (c->cn_part.n < npart) {
    c->cn_part.release();
    HIP_TRY(c->cn_part.alloc(npart));
  }
  return 0;
}

// One clh1rat from a form (numerator num[0..dimn], first-factor sign dfac,
// denominator roots[0..dimd-1]): the scaling of the gen_* loops (numerator
// by nscale tunit^k, the denominator slopes by tunit dt), the dt -> 0
// limits, and the element Schur tables of its distinct factors 1 + (cd1/cd0)T
// (h1rat_gen skips factors with |cd1| <= thresh).
int etd_set(td1c_ctx* c, int r, double dfac, const double* num, int dimn, const std::complex<double>* root, int dimd,
            double nscale, double dt, double thresh) {
  Setup& S = c->S;
  const std::complex<double> tunit = S.icomp == 0 ? std::complex<double>(-1.0, 0.0) : std::complex<double>(0.0, -1.0);
  td1c_ctx::EtdRat& R = c->etd[r];
  std::complex<double> ttmp = nscale, cd0[3], cd1[3];
  for (int k = 0; k <= dimn; ++k) {
    R.cn[k] = d2(num[k] * ttmp);
    ttmp *= tunit;
  }
  const std::complex<double> tfac = tunit * dt;
  cd1[0] = dfac * tfac;
  cd0[0] = -dfac * root[0];
  for (int i = 1; i < dimd; ++i) {
    cd1[i] = tfac;
    cd0[i] = -root[i];
  }
  R.dt = dt;
  R.dimn = dimn;
  R.dimd = dimd;
  R.rk = make_double2(1.0, 0.0);
  std::complex<double> den = 1.0;
  for (int i = 0; i < dimd; ++i) {
    R.cd0[i] = d2(cd0[i]);
    R.cd1[i] = d2(cd1[i]);
    den *= cd0[i];
  }
  const std::complex<double> nl = num[0] * std::complex<double>(nscale);
  R.numer_limit = d2(nl);
  R.denom_limit = d2(den);
  R.coeff_limit = d2(nl / den);
  auto up = [](DevBuf<double2>& b, const std::vector<double>& v) -> hipError_t {
    hipError_t e = b.alloc(std::max<size_t>(v.size() / 2, 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
  };
  for (int i = 0; i < 3; ++i) {
    R.tab[i] = -1;
    if (i >= dimd || !(std::abs(cd1[i]) > thresh)) continue;
    const std::complex<double> f = cd1[i] / cd0[i];
    for (size_t t = 0; t < c->etd_tabs.size(); ++t)
      if (c->etd_tabs[t]->fre == f.real() && c->etd_tabs[t]->fim == f.imag()) R.tab[i] = (int)t;
    if (R.tab[i] >= 0) continue;
    CnTables T;
    std::string err;
    if (!build_cn(S, c->v2jfc_h.empty() ? nullptr : c->v2jfc_h.data(), f.real(), f.imag(), T, err))
      return fail(-5, err);
    auto tab = std::make_unique<td1c_ctx::EtdTab>();
    HIP_TRY(up(tab->G, T.G));
    HIP_TRY(up(tab->Wl, T.Wl));
    HIP_TRY(up(tab->Wr, T.Wr));
    HIP_TRY(up(tab->pa, T.pa));
    HIP_TRY(up(tab->pg, T.pg));
    HIP_TRY(up(tab->binv, T.binv));
    tab->K = T.K;
    tab->fre = f.real();
    tab->fim = f.imag();
    R.tab[i] = (int)c->etd_tabs.size();
    c->etd_tabs.push_back(std::move(tab));
  }
  return 0;
}

// The inverse of clexpab's 4 x 4 Taylor matrix A(i, j) = (-i)^j / j!
// (A(i, 0) = 1, A(0, j >= 1) = 0; prop_general, src/prop/clexpab.cpp:
// 407-425), column-major, by Gauss-Jordan elimination with row pivoting.
void xab_inverse(double* out) {
  constexpr int n = 4;
  double a[n][n], b[n][n];  // [row][column]
  double fact = 1.0;
  for (int i = 0; i < n; ++i) {
    a[i][0] = 1.0;
    for (int j = 0; j < n; ++j) b[i][j] = i == j ? 1.0 : 0.0;
  }
  for (int j = 1; j < n; ++j) {
    fact *= j;
    a[0][j] = 0.0;
    for (int i = 1; i < n; ++i) a[i][j] = std::pow(-(double)i, j) / fact;
  }
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
    std::swap(a[col], a[piv]);
    std::swap(b[col], b[piv]);
    const double d = a[col][col];
    for (int j = 0; j < n; ++j) {
      a[col][j] /= d;
      b[col][j] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      const double f = a[r][col];
      for (int j = 0; j < n; ++j) {
        a[r][j] -= f * a[col][j];
        b[r][j] -= f * b[col][j];
      }
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) out[n * j + i] = b[i][j];
}

// scheme 1: cletdrb::gen (the [2/3] forms, numerators scaled by dt1 for
// both step sizes); scheme 2: cletdrk4::gen (:24-60) with gen_exp,
// gen_gfun, gen_f0fun, gen_f1fun, gen_f2fun (:134-933), the [2/2] forms.
int etd_generate(td1c_ctx* c, int scheme, double dt, int rosen, int maxcyc, double thresh) {
  Setup& S = c->S;
  const char* name = scheme == 1 ? "etdrb" : (scheme == 2 ? "etdrk4" : (scheme == 3 ? "lawrk" : "expab"));
  if (S.isplit != 1) return fail(-1, std::string(name) + " needs split_type = 1 (the split hprod)");
  if (S.lconst != 0) return fail(-1, std::string(name) + " with lconst is not provided");
  if (S.nfcore > 0 && !c->wfn0_set) return fail(-4, "frozen core needs td1c_set_wfn0 first");
  if (!(dt != 0.0) || !std::isfinite(dt)) return fail(-1, std::string(name) + " needs a finite nonzero dtime");
  if (maxcyc < 0 || !(thresh >= 0.0)) return fail(-1, "bad maxcyc or thresh");
  if (c->etd_buf.p) HIP_TRY(hipDeviceSynchronize());  // earlier steps may still read the old tables
  for (auto& t : c->etd_tabs) {
    DevBuf<double2>* e[] = {&t->G, &t->Wl, &t->Wr, &t->pa, &t->pg, &t->binv};
    for (auto* b : e) b->release();
  }
  c->etd_tabs.clear();
  c->etd_ready = false;
  if (scheme == 1) {
    const int fam[kEtdRats] = {0, 0, 1, 1, 2, 2, 3};
    const double dts[kEtdRats] = {dt, 0.5 * dt, dt, 0.5 * dt, dt, 0.5 * dt, dt};
    for (int r = 0; r < kEtdRats; ++r) {
      double dfac, num[3];
      std::complex<double> root[3];
      etd_form(fam[r], dfac, num, root);
      if (int rc = etd_set(c, r, dfac, num, 2, root, 3, dt, dts[r], thresh)) return rc;
    }
  } else if (scheme == 4) {
    // clexpab::gen (:25-110) with gen_phi4's [1/2] form (:1148-1354): numerator
    // 1 - x scaled by dt, first factor 6, roots 2 -+ sqrt(2) i
    const double num[2] = {1.0, -1.0};
    const std::complex<double> root[2] = {std::complex<double>(2.0, -1.4142135623730950488),
                                          std::complex<double>(2.0, 1.4142135623730950488)};
    if (int rc = etd_set(c, 0, 6.0, num, 1, root, 2, dt, dt, thresh)) return rc;
    xab_inverse(c->xab_ainv);
    c->xab_nstep = 0;
  } else if (scheme == 3) {
    // cllawrk::gen: dPade2p.gen(dt/2, alpha 2) is dpade_gen's 1 + (i dt/4) T,
    // the half step's factor for this dt
    if (int rc = cn_prepare(c, dt)) return rc;
  } else {
    const double dt2 = 0.5 * dt, s3 = std::sqrt(3.0);
    using cd = std::complex<double>;
    const double n_exp[3] = {12.0, 6.0, 1.0}, n_g[3] = {60.0, 6.0, 1.0}, n_f0[3] = {34860.0, 19620.0, 2657.0},
                 n_f1[3] = {2940.0, 450.0, 41.0}, n_f2[3] = {1260.0, -420.0, -13.0};
    const cd r_exp[2] = {cd(3.0, -s3), cd(3.0, s3)}, r_g[2] = {cd(4.0, -2.0), cd(4.0, 2.0)},
             r_f0[2] = {cd(762.0, -std::sqrt(189762.0)) / 221.0, cd(762.0, std::sqrt(189762.0)) / 221.0},
             r_f1[2] = {cd(51.0, -std::sqrt(633.0)) / 11.0, cd(51.0, std::sqrt(633.0)) / 11.0},
             r_f2[2] = {cd(21.0, -3.0 * std::sqrt(21.0)) / 5.0, cd(21.0, 3.0 * std::sqrt(21.0)) / 5.0};
    int rc = etd_set(c, kExp1, 1.0, n_exp, 2, r_exp, 2, 1.0, dt, thresh);
    if (!rc) rc = etd_set(c, kExp2, 1.0, n_exp, 2, r_exp, 2, 1.0, dt2, thresh);
    if (!rc) rc = etd_set(c, kGfun, 3.0, n_g, 2, r_g, 2, dt2, dt2, thresh);
    if (!rc) rc = etd_set(c, kF0, 13260.0, n_f0, 2, r_f0, 2, dt, dt, thresh);
    if (!rc) rc = etd_set(c, kF1, 330.0, n_f1, 2, r_f1, 2, dt, dt, thresh);
    if (!rc) rc = etd_set(c, kF2, 300.0, n_f2, 2, r_f2, 2, dt, dt, thresh);
    if (rc) return rc;
    c->etd[kGfun].rk = make_double2(dt2, 0.0);
    c->etd[kF0].rk = make_double2(dt / 6.0, 0.0);
    c->etd[kF1].rk = make_double2(dt / 3.0, 0.0);
    c->etd[kF2].rk = make_double2(dt / 6.0, 0.0);
  }
  if (int rc = cn_buffers(c)) return rc;
  const size_t so = (size_t)c->nbas * S.nfun, sz = so + S.lcic;
  const size_t ndn = (size_t)c->nbas * std::max(S.nfun - S.nfcore, 1);
  const size_t need = 13 * sz + so + 3 * ndn;
  if (c->etd_buf.n < need) {
    c->etd_buf.release();
    HIP_TRY(c->etd_buf.alloc(need));
  }
  if (!c->etd_lf.p) HIP_TRY(c->etd_lf.alloc(54));
  if (!c->etd_zero.p) {
    HIP_TRY(c->etd_zero.alloc(1));
    HIP_TRY(hipMemset(c->etd_zero.p, 0, sizeof(int)));
  }
  if (!c->clock.p) HIP_TRY(c->clock.alloc(1));
  c->etd_dt = dt;
  c->etd_rosen = rosen ? 1 : 0;
  c->etd_maxcyc = maxcyc;
  c->etd_thresh = thresh;
  c->etd_scheme = scheme;
  c->etd_ready = true;
  return 0;
}

// One step of the generated scheme on wfn with the field arrays d_lf[6][9]
// on the device -- cletdrb::prop4_orb (src/prop/cletdrb.cpp:321-434) or
// cletdrk4::prop (src/prop/cletdrk4.cpp:69-114) -- then the mask of
// tdse_prop12 (src/main/tdse.cpp:268-309).
int enqueue_etd_step(td1c_ctx* c, const double* d_lf, double2* wfn, int apply_mask, hipStream_t s) {
  Setup& S = c->S;
  const size_t nb = c->nbas, so = nb * S.nfun, sz = so + S.lcic, dof = nb * S.nfcore;
  const int nd = S.nfun - S.nfcore;
  const size_t ndn = nb * nd;
  double2* base = c->etd_buf.p;
  double2 *W0 = base, *W1 = base + sz, *W2 = base + 2 * sz, *W3 = base + 3 * sz, *hW0 = base + 4 * sz,
          *hW1 = base + 5 * sz, *hW2 = base + 6 * sz, *hW3 = base + 7 * sz, *tW = base + 8 * sz, *dW = base + 9 * sz;
  // an orbital block after the 13 state slots (10 of the RK-like schemes, 3K +
  // 1 of clexpab); the rational functions work on its dynamical rows
  double2* orb = base + 13 * sz;
  double2* tb[2] = {orb + so, orb + so + ndn};
  double2* vs = orb + so + 2 * ndn;
  const double *lf0 = d_lf, *lf1 = d_lf + 9, *lf3 = d_lf + 18, *dlf0 = d_lf + 27, *dv1 = d_lf + 36, *dv3 = d_lf + 45;
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  const double2 tunit = S.icomp == 0 ? make_double2(-1.0, 0.0) : make_double2(0.0, -1.0);
  // the field of the rational functions: cletdrb's timeh = t (etd_time_of_h
  // 0), cletdrk4's thalf = t + dt/2, clexpab's t
  const double* lfr = c->etd_scheme == 2 ? lf1 : lf0;
  const bool ob_el = S.n == 20;
  auto cpy = [&](double2* z, const double2* x, size_t n) { launch_zaxpby(n, one, x, zero, nullptr, z, s); };
  auto clr = [&](double2* z, size_t n) { launch_zaxpby(n, zero, nullptr, zero, nullptr, z, s); };
  // y += a x
  auto axpy = [&](double2* y, double2 a, const double2* x, size_t n) { launch_zaxpby(n, one, y, a, x, y, s); };
  auto ob_args = [&](const double* lf) {
    OneBodyElArgs ob = onebody_el_args(c, lf);
    ob.nfun = nd;
    ob.nfcore = 0;
    ob.fshift = S.nfcore;
    ob.isplit = S.isplit;
    return ob;
  };
  // out = cv (T + V(lf)) in, or cv V(lf) in (vonly), over the dynamical rows
  auto h1 = [&](const double2* in, double2* out, double2 cv, const double* lf, bool vonly) {
    CnArgs a = cn_args(c, lf, wfn);
    a.x = const_cast<double2*>(in);
    a.y = out;
    a.zfac1 = cv;
    a.flag = c->etd_zero.p;
    const OneBodyElArgs ob = ob_args(lf);
    launch_cn_stage(a, ob_el ? &ob : nullptr, c->ob_shape, c->ob_lc, vonly ? 4 : 3, s);
  };
  // the numerator sum_k cn[k] (dt h1)^k of R on in (dynamical rows) into out
  auto numer = [&](const td1c_ctx::EtdRat& R, const double2* in, double2* out) {
    launch_zaxpby(ndn, R.cn[0], in, zero, nullptr, out, s);
    const double2* src = in;
    for (int k = 1; k <= R.dimn; ++k) {
      double2* t = tb[k & 1];
      h1(src, t, make_double2(R.dt, 0.0), lfr, false);
      axpy(out, R.cn[k], t, ndn);
      src = t;
    }
  };
  // the denominator factors of R on the dynamical rows of orb, in place
  int series_err = 0;  // a failed read-back of a series flag (enqueue_series)
  auto denom = [&](const td1c_ctx::EtdRat& R) {
    for (int i = 0; i < R.dimd; ++i) {
      if (R.tab[i] < 0) continue;
      const td1c_ctx::EtdTab& T = *c->etd_tabs[R.tab[i]];
      CnArgs a = cn_args(c, lfr, orb);
      a.G = T.G.p; a.Wl = T.Wl.p; a.Wr = T.Wr.p; a.pa = T.pa.p; a.pg = T.pg.p; a.binv = T.binv.p; a.K = T.K;
      a.gstride = S.nfe;
      const std::complex<double> c0(R.cd0[i].x, R.cd0[i].y), c1(R.cd1[i].x, R.cd1[i].y);
      a.zfac0 = d2(1.0 / c0);
      a.zfac1 = d2(-(c1 / c0));
      a.thresh = c->etd_thresh;
      OneBodyElArgs ob = ob_args(lfr);
      ob.cv = a.zfac1;
      ob.stop = a.flag;
      const OneBodyElArgs* obp = ob_el ? &ob : nullptr;
      launch_cn_stage(a, obp, c->ob_shape, c->ob_lc, 2, s);
      if (S.isplit == 1 && S.icomp == 1)
        if (int rc = enqueue_series(c, a, obp, c->etd_maxcyc, s)) series_err = rc;
    }
  };
  auto cmulh = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
  // clh1rat::prod_numer: wout.ci += zfac numer_limit win.ci, wout.orb += zfac N win.orb
  auto prod_numer = [&](const td1c_ctx::EtdRat& R, double2 zfac, const double2* win, double2* wout) {
    axpy(wout + so, cmulh(zfac, R.numer_limit), win + so, S.lcic);
    numer(R, win + dof, orb + dof);
    axpy(wout + dof, zfac, orb + dof, ndn);
  };
  // clh1rat::prod_denom (with_numer false) / prod2: wout.ci += c win.ci,
  // wout.orb += zfac R(h1) win.orb
  auto prod = [&](const td1c_ctx::EtdRat& R, bool with_numer, double2 zfac, const double2* win, double2* wout) {
    const std::complex<double> z(zfac.x, zfac.y);
    const std::complex<double> cc = with_numer ? z * std::complex<double>(R.coeff_limit.x, R.coeff_limit.y)
                                               : z / std::complex<double>(R.denom_limit.x, R.denom_limit.y);
    axpy(wout + so, d2(cc), win + so, S.lcic);
    if (with_numer)
      numer(R, win + dof, orb + dof);
    else
      cpy(orb + dof, win + dof, ndn);
    denom(R);
    axpy(wout + dof, zfac, orb + dof, ndn);
  };
  auto htot = [&](const double* lf, const double2* in, double2* out) -> int {
    CiStage cs;
    cs.kout = out + so;
    if (int rc = enqueue_htotx(c, 1.0, lf, in, in + so, cs, s)) return rc;
    DtorbArgs d = dtorb_args(c, 1.0);
    d.kout = out;
    { TIMED("dtorb"); launch_dtorb(d, s); }
    HIP_TRY(join_island(c, s));
    return 0;
  };
  const auto& R = c->etd;
  if (c->etd_scheme == 4) {
    TIMED("expab");
    // clexpab::prop_general (:370-496) with K = 4; d_lf: lfield(t), then
    // lfield(timeP[k]) - lfield(t) for k = 1..3
    constexpr int K = 4, K1 = K - 1;
    double2 *P[K], *hP[K], *B[K];
    for (int k = 0; k < K; ++k) {
      P[k] = base + (size_t)k * sz;
      hP[k] = base + (size_t)(K + k) * sz;
      B[k] = base + (size_t)(2 * K + k) * sz;
    }
    double2* tw = base + (size_t)3 * K * sz;
    if (c->xab_nstep == 0) {
      cpy(P[0], wfn, sz);
      if (int rc = htot(lf0, wfn, hP[0])) return rc;
      for (int k = 1; k < K; ++k) {
        cpy(P[k], P[0], sz);
        cpy(hP[k], hP[0], sz);
      }
    } else {
      for (int k = K1; k > 0; --k) {
        cpy(P[k], P[k - 1], sz);
        cpy(hP[k], hP[k - 1], sz);
      }
      cpy(P[0], wfn, sz);
      if (int rc = htot(lf0, wfn, hP[0])) return rc;
    }
    c->xab_nstep++;
    for (int k = 0; k < K; ++k) cpy(B[k], hP[k], sz);
    for (int k = 1; k < K; ++k) {  // + tunit (V(timeP[k]) - V(t)) u_k
      h1(P[k] + dof, vs, tunit, d_lf + 9 * k, true);
      axpy(B[k] + dof, one, vs, ndn);
    }
    clr(tw, sz);
    h1(wfn + dof, tw + dof, one, lf0, false);
    launch_zaxpby(sz, tunit, tw, one, hP[0], tw, s);
    const double dt = c->etd_dt;
    int jfact = 1;
    for (int js = 1; js < K; ++js) {
      jfact *= js;
      axpy(wfn, make_double2(dt / jfact, 0.0), tw, sz);
      cpy(P[K1], tw, sz);
      clr(tw, sz);
      h1(P[K1] + dof, tw + dof, cmulh(tunit, make_double2(dt, 0.0)), lf0, false);
      for (int k = 0; k < K; ++k) axpy(tw, make_double2(c->xab_ainv[K * k + js], 0.0), B[k], sz);
    }
    jfact *= K;
    launch_zaxpby(S.lcic, make_double2(dt / jfact, 0.0), tw + so, zero, nullptr, tw + so, s);
    // expab_phi.prod on tw's dynamical orbitals, then wfn += tw
    numer(R[0], tw + dof, orb + dof);
    denom(R[0]);
    axpy(wfn + so, one, tw + so, S.lcic);
    axpy(wfn + dof, one, orb + dof, ndn);
    if (apply_mask) launch_mask(wfn, c->mask.p, S.nr, S.L1, S.nfcore, S.nfun, s);
    HIP_TRY(hipGetLastError());
    return series_err;
  }
  if (c->etd_scheme == 3) {
    TIMED("lawrk");
    // cllawrk::expwfn11 (:296-324): out = in, dpade_prod on out's dynamical
    // orbitals with the field at t + dt/2 -- x0 = (1 + i T dt/4)^-1 out, then
    // the Neumann series in V (dpade_prod.F90) -- and out.orb = -in.orb + 2 out.orb
    auto expwfn = [&](const double2* in, double2* out) {
      cpy(out, in, sz);
      CnArgs a = cn_args(c, lf1, out);
      a.zfac0 = one;
      a.thresh = c->etd_thresh;
      OneBodyElArgs ob = ob_args(lf1);
      ob.cv = a.zfac1;
      ob.stop = a.flag;
      const OneBodyElArgs* obp = ob_el ? &ob : nullptr;
      launch_cn_stage(a, obp, c->ob_shape, c->ob_lc, 2, s);
      if (int rc = enqueue_series(c, a, obp, c->etd_maxcyc, s)) series_err = rc;
      launch_zaxpby(so, make_double2(-1.0, 0.0), in, make_double2(2.0, 0.0), out, out, s);
    };
    const double dt = c->etd_dt;
    const double2 cdt1 = make_double2(dt, 0.0), cdt2 = make_double2(dt / 2.0, 0.0),
                  cdt3 = make_double2(dt / 3.0, 0.0), cdt6 = make_double2(dt / 6.0, 0.0);
    // cllawrk::prop4 (:188-280): W1 = e u0, hW1 = e N0 (e = exp(-i h1 dt/2))
    cpy(W0, wfn, sz);
    if (int rc = htot(lf0, W0, hW0)) return rc;
    expwfn(W0, W1);
    expwfn(hW0, hW1);
    launch_zaxpby(sz, cdt6, hW1, one, W1, wfn, s);
    launch_zaxpby(sz, cdt2, hW1, one, W1, W0, s);
    if (int rc = htot(lf1, W0, hW0)) return rc;
    axpy(wfn, cdt3, hW0, sz);
    launch_zaxpby(sz, cdt2, hW0, one, W1, W0, s);
    if (int rc = htot(lf1, W0, hW0)) return rc;
    axpy(wfn, cdt3, hW0, sz);
    axpy(W1, cdt1, hW0, sz);
    expwfn(W1, W0);
    if (int rc = htot(lf3, W0, hW0)) return rc;
    expwfn(wfn, W1);
    launch_zaxpby(sz, cdt6, hW0, one, W1, wfn, s);
    if (apply_mask) launch_mask(wfn, c->mask.p, S.nr, S.L1, S.nfcore, S.nfun, s);
    HIP_TRY(hipGetLastError());
    return series_err;
  }
  if (c->etd_scheme == 2) {
    TIMED("etdrk4");
    // cletdrk4::prod1 (:116-123): wout = win, its CI times rk, its
    // dynamical orbitals through R (clh1rat::prod, clh1rat.cpp:49-65)
    auto prod1 = [&](const td1c_ctx::EtdRat& Q, const double2* win, double2* wout) {
      cpy(wout, win, dof);
      launch_zaxpby(S.lcic, Q.rk, win + so, zero, nullptr, wout + so, s);
      numer(Q, win + dof, orb + dof);
      denom(Q);
      cpy(wout + dof, orb + dof, ndn);
    };
    // cletdrk4::prod2 (:125-132): wout += win with its CI times rk and its
    // dynamical orbitals through R (win's frozen-core rows are zero)
    auto prod2 = [&](const td1c_ctx::EtdRat& Q, const double2* win, double2* wout) {
      launch_zaxpby(S.lcic, Q.rk, win + so, one, wout + so, wout + so, s);
      numer(Q, win + dof, orb + dof);
      denom(Q);
      launch_zaxpby(ndn, one, orb + dof, one, wout + dof, wout + dof, s);
    };
    cpy(W0, wfn, sz);
    if (int rc = htot(lf0, W0, hW0)) return rc;
    prod1(R[kGfun], hW0, W1);
    prod1(R[kExp2], W0, tW);
    launch_zaxpby(sz, one, tW, one, W1, W1, s);
    if (int rc = htot(lf1, W1, hW1)) return rc;
    prod1(R[kGfun], hW1, W2);
    launch_zaxpby(sz, one, tW, one, W2, W2, s);
    if (int rc = htot(lf1, W2, hW2)) return rc;
    launch_zaxpby(sz, make_double2(2.0, 0.0), hW2, make_double2(-1.0, 0.0), hW0, tW, s);
    prod1(R[kGfun], tW, W3);
    prod1(R[kExp2], W1, tW);
    launch_zaxpby(sz, one, tW, one, W3, W3, s);
    if (int rc = htot(lf3, W3, hW3)) return rc;
    launch_zaxpby(sz, one, hW1, one, hW2, hW2, s);
    prod1(R[kExp1], W0, wfn);
    prod2(R[kF0], hW0, wfn);
    prod2(R[kF1], hW2, wfn);
    prod2(R[kF2], hW3, wfn);
    if (apply_mask) launch_mask(wfn, c->mask.p, S.nr, S.L1, S.nfcore, S.nfun, s);
    HIP_TRY(hipGetLastError());
    return series_err;
  }
  TIMED("etdrb");
  // stage 0: tW = -i h1 u0 + W0; u1, u3 from the phi1 numerators (and the
  // Rosenbrock term -i dt dV/dt u0)
  cpy(W0, wfn, sz);
  clr(W1, sz);
  clr(W3, sz);
  clr(tW, sz);
  h1(W0 + dof, tW + dof, one, lf0, false);
  if (int rc = htot(lf0, W0, hW0)) return rc;
  launch_zaxpby(sz, tunit, tW, one, hW0, tW, s);
  prod_numer(R[kPhi1Dt2], make_double2(0.5, 0.0), tW, W1);
  prod_numer(R[kPhi1Dt1], one, tW, W3);
  if (c->etd_rosen) {
    clr(dW, sz);
    h1(W0 + dof, dW + dof, one, dlf0, true);
    launch_zaxpby(sz, cmulh(tunit, make_double2(c->etd_dt, 0.0)), dW, zero, nullptr, dW, s);
    prod_numer(R[kPhid1Dt2], make_double2(0.25, 0.0), dW, W1);
    prod_numer(R[kPhid1Dt1], one, dW, W3);
  }
  // stage 1
  cpy(tW, W1, sz);
  cpy(W1, W0, sz);
  prod(R[kPhi1Dt2], false, one, tW, W1);
  if (int rc = htot(lf1, W1, hW1)) return rc;
  axpy(hW1, make_double2(-1.0, 0.0), hW0, sz);
  if (c->etd_rosen) {
    h1(W1 + dof, vs, tunit, dv1, true);
    axpy(hW1 + dof, one, vs, ndn);
    axpy(hW1, make_double2(-0.5, 0.0), dW, sz);
  }
  // stage 2
  cpy(W2, W1, sz);
  cpy(tW, hW1, sz);
  prod(R[kPhi2Dt2], true, one, tW, W2);
  if (int rc = htot(lf1, W2, hW2)) return rc;
  axpy(hW2, make_double2(-1.0, 0.0), hW0, sz);
  if (c->etd_rosen) {
    h1(W2 + dof, vs, tunit, dv1, true);
    axpy(hW2 + dof, one, vs, ndn);
    axpy(hW2, make_double2(-0.5, 0.0), dW, sz);
  }
  // stage 3
  cpy(tW, W3, sz);
  cpy(W3, W0, sz);
  prod(R[kPhi1Dt1], false, one, tW, W3);
  cpy(tW, hW2, sz);
  prod(R[kPhi2Dt1], true, make_double2(2.0, 0.0), tW, W3);
  if (int rc = htot(lf3, W3, hW3)) return rc;
  axpy(hW3, make_double2(-1.0, 0.0), hW0, sz);
  if (c->etd_rosen) {
    h1(W3 + dof, vs, tunit, dv3, true);
    axpy(hW3 + dof, one, vs, ndn);
    axpy(hW3, make_double2(-1.0, 0.0), dW, sz);
  }
  // summation
  cpy(wfn, W3, sz);
  launch_zaxpby(sz, make_double2(2.0, 0.0), hW1, make_double2(-1.0, 0.0), hW3, tW, s);
  prod(R[kPhi2Dt1], true, one, tW, wfn);
  launch_zaxpby(sz, one, hW1, one, hW2, tW, s);
  launch_zaxpby(sz, one, tW, make_double2(-1.0, 0.0), hW3, tW, s);
  prod(R[kPhi3Dt1], true, make_double2(-4.0, 0.0), tW, wfn);
  if (apply_mask) launch_mask(wfn, c->mask.p, S.nr, S.L1, S.nfcore, S.nfun, s);
  HIP_TRY(hipGetLastError());
  return series_err;
}

}  // namespace

int td1c_etdrb_gen(td1c_ctx* c, double dtime, int rosenbrock, int maxcyc, double thresh) {
  if (!c) return fail(-1, "null argument");
  return etd_generate(c, 1, dtime, rosenbrock, maxcyc, thresh);
}

int td1c_etdrk4_gen(td1c_ctx* c, double dtime, int maxcyc, double thresh) {
  if (!c) return fail(-1, "null argument");
  return etd_generate(c, 2, dtime, 0, maxcyc, thresh);
}

int td1c_expab_gen(td1c_ctx* c, double dtime, int maxcyc, double thresh) {
  if (!c) return fail(-1, "null argument");
  return etd_generate(c, 4, dtime, 0, maxcyc, thresh);
}

// clexpab's history times for the step at t: prop_general's shift
// (src/prop/clexpab.cpp:378-397), every slot t on the first step
static void xab_advance_times(td1c_ctx* c, double t) {
  if (c->xab_nstep == 0) {
    for (int k = 0; k < 4; ++k) c->xab_times[k] = t;
  } else {
    for (int k = 3; k > 0; --k) c->xab_times[k] = c->xab_times[k - 1];
    c->xab_times[0] = t;
  }
}

int td1c_expab_prop(td1c_ctx* c, double t, const double* lfields, double* d_wfn, int apply_mask, void* stream) {
  if (!c || !lfields || !d_wfn) return fail(-1, "null argument");
  if (!c->etd_ready || c->etd_scheme != 4) return fail(-1, "call td1c_expab_gen first");
  if (active_recorder()) return fail(-1, "exponential steps are not batched");
  double lf[36];
  for (int i = 0; i < 9; ++i) lf[i] = lfields[i];
  for (int k = 1; k < 4; ++k)
    for (int i = 0; i < 9; ++i) lf[9 * k + i] = lfields[9 * k + i] - lfields[i];
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(c->etd_lf.p, lf, 36 * sizeof(double), hipMemcpyHostToDevice, s));
  // the copy must leave lf before it goes out of scope
  HIP_TRY(hipStreamSynchronize(s));
  xab_advance_times(c, t);
  return enqueue_etd_step(c, c->etd_lf.p, reinterpret_cast<double2*>(d_wfn), apply_mask, s);
}

int td1c_expab_propagate(td1c_ctx* c, const td1c_field* field, double t0, double dtime, int nstep, double* d_wfn,
                         int apply_mask, void* stream) {
  if (!c || !field || !d_wfn) return fail(-1, "null argument");
  if (!c->etd_ready || c->etd_scheme != 4) return fail(-1, "call td1c_expab_gen first");
  if (dtime != c->etd_dt) return fail(-1, "dtime differs from the one the tables were generated for");
  if (check_field(field)) return -1;
  if (nstep < 0) return fail(-1, "nstep must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const Pulse p = make_pulse(field);
  for (int i = 0; i < nstep; ++i) {
    const double t = t0 + i * dtime;
    xab_advance_times(c, t);
    launch_xab_fields(p, t, c->xab_times, c->etd_lf.p, s);
    if (int rc = enqueue_etd_step(c, c->etd_lf.p, reinterpret_cast<double2*>(d_wfn), apply_mask, s)) return rc;
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_lawrk_gen(td1c_ctx* c, double dtime, int maxcyc, double thresh) {
  if (!c) return fail(-1, "null argument");
  return etd_generate(c, 3, dtime, 0, maxcyc, thresh);
}

static const char* etd_gen_name(int scheme) {
  return scheme == 1 ? "call td1c_etdrb_gen first" : (scheme == 2 ? "call td1c_etdrk4_gen first" : "call td1c_lawrk_gen first");
}

static int etd_prop(td1c_ctx* c, int scheme, const double* lfields, double* d_wfn, int apply_mask, void* stream) {
  if (!c || !lfields || !d_wfn) return fail(-1, "null argument");
  if (!c->etd_ready || c->etd_scheme != scheme) return fail(-1, etd_gen_name(scheme));
  if (active_recorder()) return fail(-1, "exponential steps are not batched");
  if (scheme == 3)  // the half-step factors, in case a split2 call of another dt replaced them
    if (int rc = cn_prepare(c, c->etd_dt)) return rc;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(c->etd_lf.p, lfields, 54 * sizeof(double), hipMemcpyHostToDevice, s));
  return enqueue_etd_step(c, c->etd_lf.p, reinterpret_cast<double2*>(d_wfn), apply_mask, s);
}

static int etd_propagate(td1c_ctx* c, int scheme, const td1c_field* field, double t0, double dtime, int nstep,
                         double* d_wfn, int apply_mask, void* stream) {
  if (!c || !field || !d_wfn) return fail(-1, "null argument");
  if (!c->etd_ready || c->etd_scheme != scheme) return fail(-1, etd_gen_name(scheme));
  if (dtime != c->etd_dt) return fail(-1, "dtime differs from the one the tables were generated for");
  if (scheme == 3)
    if (int rc = cn_prepare(c, c->etd_dt)) return rc;
  if (check_field(field)) return -1;
  if (nstep < 0) return fail(-1, "nstep must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const Pulse p = make_pulse(field);
  launch_clock_set(c->clock.p, t0, dtime, s);
  for (int i = 0; i < nstep; ++i) {
    launch_etd_fields(p, c->clock.p, c->etd_lf.p, s);
    if (int rc = enqueue_etd_step(c, c->etd_lf.p, reinterpret_cast<double2*>(d_wfn), apply_mask, s)) return rc;
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_etdrb_prop(td1c_ctx* c, const double* lfields, double* d_wfn, int apply_mask, void* stream) {
  return etd_prop(c, 1, lfields, d_wfn, apply_mask, stream);
}

int td1c_etdrk4_prop(td1c_ctx* c, const double* lfields, double* d_wfn, int apply_mask, void* stream) {
  return etd_prop(c, 2, lfields, d_wfn, apply_mask, stream);
}

int td1c_lawrk_prop(td1c_ctx* c, const double* lfields, double* d_wfn, int apply_mask, void* stream) {
  return etd_prop(c, 3, lfields, d_wfn, apply_mask, stream);
}

int td1c_lawrk_propagate(td1c_ctx* c, const td1c_field* field, double t0, double dtime, int nstep, double* d_wfn,
                         int apply_mask, void* stream) {
  return etd_propagate(c, 3, field, t0, dtime, nstep, d_wfn, apply_mask, stream);
}

int td1c_etdrb_propagate(td1c_ctx* c, const td1c_field* field, double t0, double dtime, int nstep, double* d_wfn,
                         int apply_mask, void* stream) {
  return etd_propagate(c, 1, field, t0, dtime, nstep, d_wfn, apply_mask, stream);
}

int td1c_etdrk4_propagate(td1c_ctx* c, const td1c_field* field, double t0, double dtime, int nstep, double* d_wfn,
                          int apply_mask, void* stream) {
  return etd_propagate(c, 2, field, t0, dtime, nstep, d_wfn, apply_mask, stream);
}

// init_prop12 (src/main/init.cpp:242-310) with the cletdrb propagator (its
// default prop_type): each cycle tests convergence by hprod_chkconv (energy
// change below tolene and out[3] below tolwfn), else takes one cletdrb step of
// dstep (icomp = 0, the field fixed at lfield: init_prop12 never advances the
// field's time) and orthonormalises (clwfn::orth).  ladapt / madapt and
// ProjHigh are identities for lconst = 0 and projhigh off.  *ene gets the
// last energy, *ncyc the last cycle index.
int td1c_init_prop12(td1c_ctx* c, double dstep, double tolene, double tolwfn, int maxcyc, const double* lfield,
                     double* d_wfn, double* ene, int* ncyc, void* stream) {
  if (!c || !lfield || !d_wfn || !ene || !ncyc) return fail(-1, "null argument");
  if (c->S.icomp != 0) return fail(-1, "imaginary-time relaxation needs a context with icomp = 0");
  if (maxcyc < 0 || !(dstep > 0.0)) return fail(-1, "maxcyc must be >= 0 and dstep > 0");
  if (int rc = etd_generate(c, 1, dstep, 1, c->cn_maxcyc, c->cn_thresh)) return rc;
  double lf[54];
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 9; ++i) lf[9 * k + i] = lfield[i];
  for (int i = 27; i < 54; ++i) lf[i] = 0.0;  // no time derivative, no differences
  double e1 = -1.0e10, chk[4] = {0, 0, 0, 0};
  int last = 0;
  for (int icyc = 0; icyc < maxcyc; ++icyc) {
    if (int rc = td1c_chkconv(c, lfield, d_wfn, chk, stream)) return rc;
    last = icyc;
    if (std::fabs(chk[0] - e1) < tolene && chk[3] < tolwfn) break;
    e1 = chk[0];
    if (int rc = td1c_etdrb_prop(c, lf, d_wfn, 0, stream)) return rc;
    if (int rc = td1c_orth(c, d_wfn, stream)) return rc;
  }
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  *ene = chk[0];
  *ncyc = last;
  return 0;
}

// ---------------------------------------------------------------------------
// Batches (scans of pulses, grids and lmax): one launch per recorded site
// for all members.  No counterpart in the reference, whose scans run one
// propagation per process.
// ---------------------------------------------------------------------------
int td1c_batch_create(td1c_ctx* const* members, int n, td1c_batch** out) {
  if (!members || !out || n <= 0) return fail(-1, "null argument or empty batch");
  *out = nullptr;
  td1c_ctx* c0 = members[0];
  if (!c0) return fail(-1, "null member");
  for (int i = 0; i < n; ++i) {
    td1c_ctx* c = members[i];
    if (!c) return fail(-1, "null member");
    for (int j = 0; j < i; ++j)
      if (members[j] == c) return fail(-1, "a context appears twice in the batch");
    const Setup& S = c->S;
    if (c->device != c0->device) return fail(-1, "batch members must share a device");
    if (S.n != 20) return fail(-1, "batched propagation needs ndvr = 20");
    if (S.nfcore != 0 || S.nact != 0)
      return fail(-1, "batched propagation holds dynamical cores only (no frozen cores, no active orbitals)");
    if (S.nfun != c0->S.nfun || c->mixed != c0->mixed || S.isplit != c0->S.isplit ||
        S.igauge != c0->S.igauge || S.icomp != c0->S.icomp)
      return fail(-1, "batch members must share nfun, precision, split_type, gauge and icomp");
  }
  td1c_batch* b = new td1c_batch();
  b->mem.assign(members, members + n);
  b->device = c0->device;
  b->tvar = c0->tvar;
  *out = b;
  return 0;
}

int td1c_batch_destroy(td1c_batch* b) {
  if (!b) return 0;
  HIP_TRY(hipDeviceSynchronize());  // replays of the batch's graph may still read its table
  delete b;
  return 0;
}

int td1c_batch_set_graph(td1c_batch* b, int on) {
  if (!b) return fail(-1, "null argument");
  b->use_graph = on != 0;
  return 0;
}

int td1c_batch_set_shapes(td1c_batch* b, int elements_per_block, int l_per_block, int transform_variant) {
  if (!b) return fail(-1, "null argument");
  const int shape = elements_per_block == 3 ? 0 : (elements_per_block == 6 ? 1 : (elements_per_block == 12 ? 2 : -1));
  if (shape < 0) return fail(-1, "elements_per_block must be 3, 6 or 12");
  if (l_per_block < 1) return fail(-1, "l_per_block must be >= 1");
  if (transform_variant != 0 && transform_variant != 1) return fail(-1, "transform_variant must be 0 or 1");
  b->ob_shape = shape;
  b->ob_lc = l_per_block;
  b->tvar = transform_variant;
  b->key.clear();  // record again
  return 0;
}

namespace {

extern "C++" {
template <class T>
void key_put(std::vector<unsigned char>& k, const T& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}
}

// Blocks [row][elem] of blk doubles gathered from tables of rows x elems
// blocks: a table joins when each of its blocks equals the block already
// there (bitwise), and fills the blocks nobody had.
struct TableUnion {
  int rows = 0, elems = 0;
  size_t blk = 0;
  std::vector<double> v;
  std::vector<char> set;
  void init(int r, int e, size_t b) {
    rows = r; elems = e; blk = b;
    v.assign((size_t)r * e * b, 0.0);
    set.assign((size_t)r * e, 0);
  }
  bool add(const double* t, int r, int e) {
    if (r > rows || e > elems) return false;
    for (int i = 0; i < r; ++i)
      for (int k = 0; k < e; ++k) {
        const size_t u = (size_t)i * elems + k;
        if (set[u] && std::memcmp(&v[u * blk], t + ((size_t)i * e + k) * blk, blk * sizeof(double)) != 0) return false;
      }
    for (int i = 0; i < r; ++i)
      for (int k = 0; k < e; ++k) {
        const size_t u = (size_t)i * elems + k;
        if (!set[u]) {
          std::memcpy(&v[u * blk], t + ((size_t)i * e + k) * blk, blk * sizeof(double));
          set[u] = 1;
        }
      }
    return true;
  }
};

extern "C++" {
template <class R, class F>
void patch_recs(std::vector<LaunchRec>& recs, int kind0, int kind1, F f) {
  for (LaunchRec& L : recs) {
    if (L.kind != kind0 && L.kind != kind1) continue;
    R r;
    std::memcpy(&r, L.args.data(), sizeof(R));
    f(r);
    std::memcpy(L.args.data(), &r, sizeof(R));
  }
}
}

// Points the members' recorded element-table arguments at shared copies
// where their own tables are equal: a scan over pulses on grids with the
// same element boundaries then reads each element block once per launch.
int batch_share_tables(td1c_batch* b, std::vector<std::vector<LaunchRec>>& recs) {
  const int nmem = (int)b->mem.size();
  int L1 = 0, L2 = 0, ne = 0;
  for (td1c_ctx* c : b->mem) {
    L1 = std::max(L1, c->S.L1);
    L2 = std::max(L2, c->S.L2);
    ne = std::max(ne, c->S.nfe);
  }
  const int ni = b->mem[0]->S.n - 1, np = b->mem[0]->S.n + 1;
  for (int& s : b->shared) s = 0;
  // Crank-Nicolson interior inverses (complex), [l][e][k][i]
  if (b->mem[0]->S.isplit == 1) {
    TableUnion u;
    u.init(L1, ne, (size_t)2 * ni * ni);
    std::vector<char> ok(nmem, 0);
    for (int m = 0; m < nmem; ++m) {
      td1c_ctx* c = b->mem[m];
      if (c->cnG_h.size() == (size_t)c->S.L1 * c->S.nfe * 2 * ni * ni) ok[m] = u.add(c->cnG_h.data(), c->S.L1, c->S.nfe);
    }
    b->u_cnG.release();
    HIP_TRY(b->u_cnG.alloc(u.v.size() / 2));
    HIP_TRY(hipMemcpy(b->u_cnG.p, u.v.data(), u.v.size() * sizeof(double), hipMemcpyHostToDevice));
    for (int m = 0; m < nmem; ++m)
      if (ok[m]) {
        ++b->shared[0];
        patch_recs<CnRec>(recs[m], kRecCnBridge, kRecCnInterior, [&](CnRec& r) {
          r.a.G = b->u_cnG.p;
          r.a.gstride = ne;
        });
      }
  }
  // Poisson interior inverses, transposed as on the device, [L][e][k][i]
  {
    TableUnion u;
    u.init(L2, ne, (size_t)ni * ni);
    std::vector<char> ok(nmem, 0);
    for (int m = 0; m < nmem; ++m) {
      const Setup& S = b->mem[m]->S;
      if (S.pG.size() != (size_t)S.L2 * S.nfe * ni * ni) continue;
      std::vector<double> gt(S.pG.size());
      for (size_t blk = 0; blk < S.pG.size() / ((size_t)ni * ni); ++blk)
        for (int i = 0; i < ni; ++i)
          for (int k = 0; k < ni; ++k) gt[blk * ni * ni + (size_t)k * ni + i] = S.pG[blk * ni * ni + (size_t)i * ni + k];
      ok[m] = u.add(gt.data(), S.L2, S.nfe);
    }
    b->u_pG.release();
    HIP_TRY(b->u_pG.upload(u.v));
    for (int m = 0; m < nmem; ++m)
      if (ok[m]) {
        ++b->shared[1];
        patch_recs<PoissonRec>(recs[m], kRecPoissonBridge, kRecPoissonInterior, [&](PoissonRec& r) {
          r.a.G = b->u_pG.p;
          r.a.gstride = ne;
        });
      }
  }
  // Poisson row tables at [L][point]: when every member shares the
  // inverses, the same element geometry gives the same numbers at the same
  // (L, point); the right coupling Wr of a member's last element (zero
  // there, the box edge, and never read) takes the value of the members
  // for which that element is an inner one, so members join largest first
  for (auto& t : b->u_prow) t.release();
  b->rstride = 0;
  b->shared[3] = 0;
  if (b->shared[1] == nmem) {
    const int n = b->mem[0]->S.n, rs = ne * n - 1;
    std::vector<int> order(nmem);
    for (int m = 0; m < nmem; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->mem[x]->S.nfe > b->mem[y]->S.nfe; });
    TableUnion u[5];
    for (auto& t : u) t.init(L2, rs, 1);
    bool all = true;
    for (int m : order) {
      const Setup& S = b->mem[m]->S;
      const std::vector<double>* tab[5] = {&S.pWl, &S.pWr, &S.d2invr, &S.d2rpl1, &S.d2fac2};
      const size_t want = (size_t)S.L2 * S.nr;
      if (S.n != n) all = false;
      for (int k = 0; k < 5 && all; ++k) {
        if (tab[k]->size() != want) {
          all = false;
          break;
        }
        std::vector<double> t(*tab[k]);
        if (k == 1)
          for (int L = 0; L < S.L2; ++L)
            for (int r = (S.nfe - 1) * n; r < S.nr; ++r) {
              const size_t q = (size_t)L * rs + r;
              if (u[1].set[q]) t[(size_t)L * S.nr + r] = u[1].v[q];
            }
        all = u[k].add(t.data(), S.L2, S.nr);
      }
      if (!all) break;
    }
    if (all) {
      for (int k = 0; k < 5; ++k) HIP_TRY(b->u_prow[k].upload(u[k].v));
      b->rstride = rs;
      b->shared[3] = nmem;
    }
  }
  // element one-body matrices, [e][k][m] each
  {
    TableUnion u0, u1, u2;
    u0.init(1, ne, (size_t)np * np);
    u1.init(1, ne, (size_t)np * np);
    u2.init(1, ne, (size_t)np * np);
    std::vector<char> ok(nmem, 0);
    for (int m = 0; m < nmem; ++m) {
      const Setup& S = b->mem[m]->S;
      const int nu = (int)(S.eK.size() / ((size_t)np * np));
      if (S.eK0.size() != S.eK.size() || S.eP.size() != S.eK.size() || nu > ne) continue;
      // the three tables join together or not at all
      TableUnion t0 = u0, t1 = u1, t2 = u2;
      if (t0.add(S.eK0.data(), 1, nu) && t1.add(S.eK.data(), 1, nu) && t2.add(S.eP.data(), 1, nu)) {
        u0 = std::move(t0);
        u1 = std::move(t1);
        u2 = std::move(t2);
        ok[m] = 1;
      }
    }
    b->u_eK0.release();
    b->u_eK.release();
    b->u_eP.release();
    HIP_TRY(b->u_eK0.upload(u0.v));
    HIP_TRY(b->u_eK.upload(u1.v));
    HIP_TRY(b->u_eP.upload(u2.v));
    for (int m = 0; m < nmem; ++m)
      if (ok[m]) {
        ++b->shared[2];
        patch_recs<OneBodyRec>(recs[m], kRecOneBody, kRecOneBody, [&](OneBodyRec& r) {
          r.a.eK0 = b->u_eK0.p;
          r.a.eK = b->u_eK.p;
          r.a.eP = b->u_eP.p;
        });
      }
  }
  return 0;
}

// Records one step of every member and lays the sites out in a device table:
// per site, the members' argument records, then start[nmem + 1], then
// dims[nmem][3], each 256-byte aligned; identical sites share one block.
int batch_record(td1c_batch* b, const std::vector<Pulse>& pulses, int rk_order, double dtime, double* const* d_wfns,
                 int apply_mask) {
  const int nmem = (int)b->mem.size();
  std::vector<std::vector<LaunchRec>> recs(nmem);
  for (int m = 0; m < nmem; ++m) {
    td1c_ctx* c = b->mem[m];
    int ns;
    StageNodes sn;
    if (int rc = step_nodes(c, rk_order, sn, ns)) return rc;
    Recorder R;
    const int tpath = c->tpath, shape = c->ob_shape, lc = c->ob_lc, tvar = c->tvar;
    const bool capturing = c->capturing;
    c->tpath = 2;  // the matrix-core transforms (their batched form)
    c->ob_shape = b->ob_shape;
    c->ob_lc = b->ob_lc;
    c->tvar = b->tvar;
    c->capturing = true;  // no per-kernel timing events
    active_recorder() = &R;
    launch_field_stages(pulses[m], c->clock.p, sn, c->lf_dev.p, c->side[1]);
    int rc = enqueue_prop_step(c, rk_order, ns, dtime, reinterpret_cast<double2*>(d_wfns[m]), apply_mask, c->side[1],
                               true);
    active_recorder() = nullptr;
    c->tpath = tpath;
    c->ob_shape = shape;
    c->ob_lc = lc;
    c->tvar = tvar;
    c->capturing = capturing;
    if (rc) return rc;
    if (!R.error.empty()) return fail(-1, "not batched: " + R.error);
    recs[m] = std::move(R.recs);
  }
  const size_t nsite = recs[0].size();
  for (int m = 1; m < nmem; ++m) {
    if (recs[m].size() != nsite) return fail(-1, "batch members record different steps");
    for (size_t i = 0; i < nsite; ++i) {
      const LaunchRec &a = recs[0][i], &r = recs[m][i];
      if (a.kind != r.kind || a.block != r.block || std::memcmp(a.var, r.var, sizeof a.var) != 0 ||
          a.args.size() != r.args.size())
        return fail(-1, "batch members record different launches");
    }
  }
  // the old tables may still be read by queued launches
  HIP_TRY(hipDeviceSynchronize());
  if (int rc = batch_share_tables(b, recs)) return rc;
  // the terms of each half-step series: the bridge pass of the first term
  // (mode 0) opens a series, each later bridge pass starts the next term; the
  // right-hand side of a later term (cn = 2) precedes its bridge pass
  if (!b->cn_word.p) {
    HIP_TRY(b->cn_word.alloc(1));
    HIP_TRY(hipMemset(b->cn_word.p, 0, sizeof(int)));
  }
  for (int m = 0; m < nmem; ++m) {
    int term = 0;
    for (LaunchRec& L : recs[m]) {
      if (L.kind == kRecCnBridge || L.kind == kRecCnInterior || L.kind == kRecCnCheck) {
        CnRec r;
        std::memcpy(&r, L.args.data(), sizeof r);
        if (L.kind == kRecCnBridge) term = r.mode == 0 ? 1 : term + 1;
        r.gate = SeriesGate{b->cn_word.p, term};
        std::memcpy(L.args.data(), &r, sizeof r);
      } else if (L.kind == kRecOneBody && L.var[1] == 2) {
        OneBodyRec r;
        std::memcpy(&r, L.args.data(), sizeof r);
        r.gate = SeriesGate{b->cn_word.p, term + 1};
        std::memcpy(L.args.data(), &r, sizeof r);
      }
    }
  }
  auto align = [](size_t x) { return (x + 255) / 256 * 256; };
  std::vector<unsigned char> host;
  std::map<std::string, size_t> seen;
  std::vector<td1c_batch::Site> sites(nsite);
  for (size_t i = 0; i < nsite; ++i) {
    td1c_batch::Site& st = sites[i];
    st.proto = recs[0][i];
    st.proto.args.clear();
    const size_t rsz = recs[0][i].args.size();
    if (rsz != rec_size(st.proto.kind)) return fail(-1, "batch record size mismatch");
    std::vector<int> start(nmem + 1, 0), dims(3 * nmem, 1);
    size_t lds = 0;
    long long tot = 0;
    for (int m = 0; m < nmem; ++m) {
      const LaunchRec& r = recs[m][i];
      start[m] = (int)tot;
      for (int d = 0; d < 3; ++d) dims[3 * m + d] = r.grid[d];
      const long long nb = (long long)r.grid[0] * r.grid[1] * r.grid[2];
      tot += nb;
      st.max_blocks = std::max(st.max_blocks, (int)std::min<long long>(nb, 0x7fffffffLL));
      lds = std::max(lds, r.lds);
    }
    if (tot <= 0 || tot > 0x7fffffffLL) return fail(-1, "batched launch too large");
    start[nmem] = (int)tot;
    st.total = (int)tot;
    st.lds = lds;
    BatchMap probe{nullptr, nullptr, nmem, 0};
    if (!launch_batch_kernels(st.proto, nullptr, probe, 0, 0, nullptr, true) &&
        !launch_batch_h1itr(st.proto, nullptr, probe, 0, 0, nullptr, true) &&
        !launch_batch_island(st.proto, nullptr, probe, 0, 0, nullptr, true))
      return fail(-1, "no batched kernel for launch kind " + std::to_string(st.proto.kind));
    std::string blob;
    for (int m = 0; m < nmem; ++m) blob.append(recs[m][i].args.begin(), recs[m][i].args.end());
    const size_t nrec = blob.size();
    blob.append(reinterpret_cast<const char*>(start.data()), start.size() * sizeof(int));
    blob.append(reinterpret_cast<const char*>(dims.data()), dims.size() * sizeof(int));
    auto it = seen.find(blob);
    size_t off;
    if (it != seen.end()) {
      off = it->second;
    } else {
      off = align(host.size());
      host.resize(off);
      host.insert(host.end(), blob.begin(), blob.begin() + nrec);
      host.resize(align(host.size()));
      host.insert(host.end(), blob.begin() + nrec, blob.begin() + nrec + start.size() * sizeof(int));
      host.resize(align(host.size()));
      host.insert(host.end(), blob.begin() + nrec + start.size() * sizeof(int), blob.end());
      seen.emplace(std::move(blob), off);
    }
    st.off_recs = off;
    st.off_start = align(off + nrec);
    st.off_dims = align(st.off_start + start.size() * sizeof(int));
  }
  // Interior passes over shared tables (the Poisson solve, the clh1itr
  // half step): items (member, pair or orbital) in groups, ordered by extent
  // (L, element chunks) so that a group's items end near the same
  // (L, chunk); jobs (group, L, chunk) for every (L, chunk) that some item of
  // the group covers.  Sites with the same items share one layout.
  struct Item { int m, p, L, nch; };
  std::map<std::vector<int>, std::array<size_t, 4>> layouts;  // items -> off_items, off_jobs, njobs, ngroups
  auto layout = [&](std::vector<Item> items, int slots) -> std::array<size_t, 4> {
    std::stable_sort(items.begin(), items.end(),
                     [](const Item& x, const Item& y) { return x.L != y.L ? x.L > y.L : x.nch > y.nch; });
    std::vector<int> key{slots};
    for (const Item& it : items) key.insert(key.end(), {it.m, it.p, it.L, it.nch});
    auto f = layouts.find(key);
    if (f != layouts.end()) return f->second;
    const int ng = ((int)items.size() + slots - 1) / slots;
    std::vector<int2> sl((size_t)ng * slots, make_int2(-1, 0));
    std::vector<int> gL(ng, 0), gC(ng, 0);
    int nL = 0, nC = 0;
    for (size_t k = 0; k < items.size(); ++k) {
      const int g = (int)k / slots;
      sl[k] = make_int2(items[k].m, items[k].p);
      gL[g] = std::max(gL[g], items[k].L);
      gC[g] = std::max(gC[g], items[k].nch);
      nL = std::max(nL, gL[g]);
      nC = std::max(nC, gC[g]);
    }
    std::vector<int> jobs;
    for (int L = 0; L < nL; ++L)
      for (int c = 0; c < nC; ++c)
        for (int g = 0; g < ng; ++g)
          if (L < gL[g] && c < gC[g]) jobs.insert(jobs.end(), {g, L, c});
    std::array<size_t, 4> out;
    out[0] = align(host.size());
    host.resize(out[0]);
    host.insert(host.end(), reinterpret_cast<const unsigned char*>(sl.data()),
                reinterpret_cast<const unsigned char*>(sl.data() + sl.size()));
    out[1] = align(host.size());
    host.resize(out[1]);
    host.insert(host.end(), reinterpret_cast<const unsigned char*>(jobs.data()),
                reinterpret_cast<const unsigned char*>(jobs.data() + jobs.size()));
    host.resize(align(host.size()));
    out[2] = jobs.size() / 3;
    out[3] = (size_t)ng;
    layouts.emplace(std::move(key), out);
    return out;
  };
  b->pgroups = b->cgroups = 0;
  for (size_t i = 0; i < nsite; ++i) {
    td1c_batch::Site& st = sites[i];
    const int kind = st.proto.kind;
    if (kind == kRecPoissonInterior && b->shared[3] == nmem && (st.proto.var[0] == 0 || st.proto.var[0] == 2)) {
      std::vector<PoissonRec> pr(nmem);
      bool ok = true;
      for (int m = 0; m < nmem && ok; ++m) {
        std::memcpy(&pr[m], recs[m][i].args.data(), sizeof(PoissonRec));
        const PoissonRec& r = pr[m];
        ok = r.a.ni == 19 && r.a.n == 20 && r.a.G == b->u_pG.p && r.epb == pr[0].epb && r.epb <= 32 &&
             r.epb * 19 <= kPoissonInteriorThreads && r.a.nfe * r.a.n - 1 == r.a.nr && r.a.nr <= b->rstride;
      }
      if (!ok) continue;
      std::vector<Item> items;
      for (int m = 0; m < nmem; ++m)
        for (int p = 0; p < pr[m].a.npair; ++p) items.push_back({m, p, pr[m].a.L2, pr[m].nchunk});
      const auto lay = layout(std::move(items), kPoissonSlots);
      if (lay[2] > 0x7fffffffULL) continue;
      st.grouped = 1;
      st.off_items = lay[0];
      st.off_jobs = lay[1];
      st.njobs = (int)lay[2];
      st.psh = PoissonShared{b->u_pG.p, b->u_prow[0].p, b->u_prow[1].p, b->u_prow[2].p, b->u_prow[3].p,
                             b->u_prow[4].p, pr[0].a.gstride, b->rstride, pr[0].epb};
      b->pgroups = (int)lay[3];
    } else if (kind == kRecCnInterior && b->shared[0] == nmem && st.proto.block == kCnInteriorThreads) {
      std::vector<CnRec> cr(nmem);
      bool ok = true;
      for (int m = 0; m < nmem && ok; ++m) {
        std::memcpy(&cr[m], recs[m][i].args.data(), sizeof(CnRec));
        const CnRec& r = cr[m];
        ok = r.a.ni == 19 && r.a.n == 20 && r.a.G == b->u_cnG.p && r.a.nfun - r.a.nfcore == 1 &&
             r.epb == cr[0].epb && r.epb <= 32 && r.epb * 19 <= kCnInteriorThreads && r.a.nb == r.a.nfe - 1 &&
             r.a.nfe <= r.a.gstride && r.a.gstride == cr[0].a.gstride;
      }
      if (!ok) continue;
      std::vector<Item> items;
      for (int m = 0; m < nmem; ++m) items.push_back({m, 0, cr[m].a.L1, cr[m].nchunk});
      const auto lay = layout(std::move(items), kCnGroup);
      if (lay[2] > 0x7fffffffULL) continue;
      st.grouped = 2;
      st.off_items = lay[0];
      st.off_jobs = lay[1];
      st.njobs = (int)lay[2];
      st.csh = CnShared{b->u_cnG.p, cr[0].a.gstride, cr[0].epb, cr[0].gate};
      b->cgroups = (int)lay[3];
    }
  }
  // the old table may still be read by queued launches
  HIP_TRY(hipDeviceSynchronize());
  if (b->gexec) {
    HIP_TRY(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  b->table.release();
  HIP_TRY(b->table.alloc(std::max<size_t>(host.size(), 256)));
  HIP_TRY(hipMemcpy(b->table.p, host.data(), host.size(), hipMemcpyHostToDevice));
  b->table_bytes = host.size();
  b->sites = std::move(sites);
  return 0;
}

hipError_t batch_launch_step(td1c_batch* b, hipStream_t s) {
  const int nmem = (int)b->mem.size();
  for (const td1c_batch::Site& st : b->sites) {
    const unsigned char* base = b->table.p;
    BatchMap map{reinterpret_cast<const int*>(base + st.off_start), reinterpret_cast<const int*>(base + st.off_dims),
                 nmem, 0};
    const void* recs = base + st.off_recs;
    if (st.grouped == 1) {
      if (!launch_poisson_interior_grouped(st.proto.var[0], recs, reinterpret_cast<const int2*>(base + st.off_items),
                                           reinterpret_cast<const int*>(base + st.off_jobs), st.njobs, st.psh, s))
        return hipErrorInvalidValue;
      continue;
    }
    if (st.grouped == 2) {
      launch_cn_interior_grouped(recs, reinterpret_cast<const int2*>(base + st.off_items),
                                 reinterpret_cast<const int*>(base + st.off_jobs), st.njobs, st.csh, s);
      continue;
    }
    int grid = st.total;
    if (batch_skips(st.proto)) {
      // two blocks of the largest member per hardware block
      map.per_member = std::max(1, (st.max_blocks + 1) / 2);
      grid = (int)std::min<long long>((long long)nmem * map.per_member, 0x7fffffffLL);
    }
    if (!launch_batch_kernels(st.proto, recs, map, grid, st.lds, s, false) &&
        !launch_batch_h1itr(st.proto, recs, map, grid, st.lds, s, false) &&
        !launch_batch_island(st.proto, recs, map, grid, st.lds, s, false))
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

int td1c_batch_propagate(td1c_batch* b, const td1c_field* fields, int rk_order, double t0, double dtime, int nstep,
                         double* const* d_wfns, int apply_mask, void* stream) {
  if (!b || !fields || !d_wfns) return fail(-1, "null argument");
  const int nmem = (int)b->mem.size();
  for (int m = 0; m < nmem; ++m) {
    if (!d_wfns[m]) return fail(-1, "null wave function");
    if (check_field(&fields[m])) return -1;
  }
  if (nstep < 0) return fail(-1, "nstep must be >= 0");
  {
    int ns;
    StageNodes sn;
    if (int rc = step_nodes(b->mem[0], rk_order, sn, ns)) return rc;
  }
  if (nstep == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  std::vector<Pulse> pulses(nmem);
  for (int m = 0; m < nmem; ++m) pulses[m] = make_pulse(&fields[m]);
  if (b->mem[0]->S.isplit == 1)
    for (td1c_ctx* c : b->mem)
      if (int rc = cn_prepare(c, dtime)) return rc;
  // the recorded step depends on the call's arguments and on the members'
  // tables that a new dt or frozen core reallocates
  std::vector<unsigned char> key;
  key_put(key, rk_order);
  key_put(key, apply_mask);
  key_put(key, dtime);
  key_put(key, b->ob_shape);
  key_put(key, b->ob_lc);
  key_put(key, b->tvar);
  for (int m = 0; m < nmem; ++m) {
    td1c_ctx* c = b->mem[m];
    key_put(key, d_wfns[m]);
    key_put(key, pulses[m]);
    key_put(key, c->cnG.p);
    key_put(key, c->cn_part.p);
    key_put(key, c->cn_dt);
  }
  if (key != b->key) {
    b->key.clear();
    if (int rc = batch_record(b, pulses, rk_order, dtime, d_wfns, apply_mask)) return rc;
    b->key = std::move(key);
  }
  for (td1c_ctx* c : b->mem) launch_clock_set(c->clock.p, t0, dtime, s);
  if (!b->use_graph) {
    for (int i = 0; i < nstep; ++i)
      if (hipError_t e = batch_launch_step(b, s)) return fail(-3, std::string("batched launch: ") + hipGetErrorString(e));
    return 0;
  }
  if (!b->gexec) {
    if (!b->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = batch_launch_step(b, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (le != hipSuccess || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(-3, std::string("batch capture: ") + hipGetErrorString(le != hipSuccess ? le : e));
    }
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      b->gexec = nullptr;
      return fail(-3, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  for (int i = 0; i < nstep; ++i) HIP_TRY(hipGraphLaunch(b->gexec, s));
  HIP_TRY(hipGetLastError());
  return 0;
}

int td1c_batch_autotune(td1c_batch* b, const td1c_field* fields, int rk_order, double dtime,
                        double* const* d_scratch, void* stream) {
  if (!b || !fields || !d_scratch) return fail(-1, "null argument");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  std::string log = "[";
  int rc = 0;
  float best = 1e30f;
  int bs = b->ob_shape, bl = b->ob_lc, bt = b->tvar;
  auto trial = [&](int shape, int lc, int tvar) -> int {
    b->ob_shape = shape;
    b->ob_lc = lc;
    b->tvar = tvar;
    b->key.clear();
    // record and warm (one step), then two timed steps, on the scratch states
    if (int r = td1c_batch_propagate(b, fields, rk_order, 0.0, dtime, 1, d_scratch, 1, stream)) return r;
    HIP_TRY(hipEventRecord(e0, s));
    if (int r = td1c_batch_propagate(b, fields, rk_order, dtime, dtime, 2, d_scratch, 1, stream)) return r;
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    ms *= 0.5f;
    static const int ge[3] = {3, 6, 12};
    char buf[96];
    std::snprintf(buf, sizeof buf, "%s[%d, %d, %d, %.4f]", log.size() > 1 ? ", " : "", ge[shape], lc, tvar, ms);
    log += buf;
    if (ms < best) {
      best = ms;
      bs = shape;
      bl = lc;
      bt = tvar;
    }
    return 0;
  };
  // the one-body shape first (transform variant fixed), then the transform variant
  const int t0 = b->tvar;
  for (int shape = 0; shape < 3 && !rc; ++shape)
    for (int lc : {2, 4, 8, 16}) {
      rc = trial(shape, lc, t0);
      if (rc) break;
    }
  if (!rc) rc = trial(bs, bl, 1 - t0);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->ob_shape = bs;
  b->ob_lc = bl;
  b->tvar = bt;
  b->key.clear();  // the next call records with its own wave functions
  b->tune_log = "{\"ms_per_step\": " + log + "]}";
  return rc;
}

int td1c_batch_get_info(const td1c_batch* b, char* buf, int buflen) {
  if (!b || !buf || buflen <= 0) return fail(-1, "null argument");
  long long blocks = 0;
  for (const auto& st : b->sites) blocks += st.grouped ? st.njobs : st.total;
  const char* ge[3] = {"3", "6", "12"};
  std::string js = "{\"members\": " + std::to_string(b->mem.size()) + ", \"sites_per_step\": " +
                   std::to_string(b->sites.size()) + ", \"blocks_per_step\": " + std::to_string(blocks) +
                   ", \"table_bytes\": " + std::to_string(b->table_bytes) + ", \"onebody\": [" + ge[b->ob_shape] +
                   ", " + std::to_string(b->ob_lc) + "], \"transform_variant\": " + std::to_string(b->tvar) +
                   ", \"graph\": " + (b->use_graph ? "true" : "false") + ", \"shared_tables\": {\"cn\": " +
                   std::to_string(b->shared[0]) + ", \"poisson\": " + std::to_string(b->shared[1]) +
                   ", \"onebody\": " + std::to_string(b->shared[2]) + ", \"poisson_rows\": " +
                   std::to_string(b->shared[3]) + "}, \"poisson_groups\": " + std::to_string(b->pgroups) +
                   ", \"cn_groups\": " + std::to_string(b->cgroups) +
                   (b->tune_log.empty() ? std::string() : ", \"tune\": " + b->tune_log) + "}";
  if ((int)js.size() + 1 > buflen) return fail(-2, "buffer too small");
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return 0;
}

int td1c_field_value(const td1c_field* f, double t, double lfield[9]) {
  if (!f || !lfield) return fail(-1, "null argument");
  for (int i = 0; i < 9; ++i) lfield[i] = 0.0;
  if (f->env_type < 0 || f->env_type > 2 || f->gauge < 0 || f->gauge > 1) return fail(-1, "bad field");
  pulse_lfield(make_pulse(f), t, lfield);
  return 0;
}

int td1c_field_der(const td1c_field* f, double t, double dfield[9]) {
  if (!f || !dfield) return fail(-1, "null argument");
  for (int i = 0; i < 9; ++i) dfield[i] = 0.0;
  if (f->env_type < 0 || f->env_type > 2 || f->gauge < 0 || f->gauge > 1) return fail(-1, "bad field");
  pulse_dlfield(make_pulse(f), t, dfield);
  return 0;
}

double td1c_field_period(const td1c_field* f) {
  if (!f) return 0.0;
  return 2.0 * kPi / (1239.84190 / (f->wlen * 27.2113845));
}

int td1c_set_timing(td1c_ctx* c, int on) {
  if (!c) return fail(-1, "null argument");
  if (on && c->ev_pool.empty()) {
    c->ev_pool.resize(8192);
    for (auto& e : c->ev_pool) HIP_TRY(hipEventCreate(&e));
  }
  HIP_TRY(hipDeviceSynchronize());
  c->timing = on != 0;
  c->recs.clear();
  c->ev_next = 0;
  c->timing_agg.clear();
  return 0;
}

int td1c_set_timing_kernel(td1c_ctx* c, const char* name) {
  if (!c) return fail(-1, "null argument");
  c->timing_only = name ? name : "";
  return 0;
}

int td1c_get_timing(td1c_ctx* c, char* buf, int buflen) {
  if (!c || !buf || buflen <= 0) return fail(-1, "null argument");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(harvest_timing(c));
  std::string out = "{";
  for (size_t k = 0; k < c->timing_agg.size(); ++k) {
    auto& a = c->timing_agg[k];
    char tmp[160];
    std::snprintf(tmp, sizeof(tmp), "%s\"%s\": [%.6f, %d]", k ? ", " : "", a.first.c_str(), a.second.first,
                  a.second.second);
    out += tmp;
  }
  out += "}";
  c->timing_agg.clear();
  if ((int)out.size() + 1 > buflen) return fail(-1, "timing buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return 0;
}

int td1c_set_graph(td1c_ctx* c, int on) {
  if (!c) return fail(-1, "null argument");
  c->use_graph = on != 0;
  return 0;
}

int td1c_set_transform_path(td1c_ctx* c, int path) {
  if (!c) return fail(-1, "null argument");
  if (path < 0 || path > 2) return fail(-1, "transform path must be 0, 1 or 2");
  if (path != c->tpath && c->gexec) {
    // a captured step holds the kernels of the old path: capture again
    HIP_TRY(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  c->tpath = path;
  return 0;
}

int td1c_bind(td1c_ctx* c) {
  g_bound = c;
  return 0;
}

void hprod_htot_(const double* dtime, const double* lfield, const double* wfn, const double* cic, double* hwfn,
                 double* hcic) {
  td1c_ctx* c = g_bound;
  if (!c) {
    std::fprintf(stderr, "hprod_htot_: no bound td1c context\n");
    std::abort();
  }
  const size_t no = (size_t)c->nbas * c->S.nfun, nc = (size_t)c->S.lcic;
  auto chk = [](hipError_t e, const char* what) {
    if (e != hipSuccess) {
      std::fprintf(stderr, "hprod_htot_: %s: %s\n", what, hipGetErrorString(e));
      std::abort();
    }
  };
  // the staging copies live as long as the context: one allocation, on the
  // first call, not one per hprod
  if (!c->legacy_stage.p) chk(c->legacy_stage.alloc(2 * (no + nc)), "hipMalloc");
  double2* dw = c->legacy_stage.p;
  double2* dh = dw + no;
  double2* dc = dh + no;
  double2* dhc = dc + nc;
  chk(hipMemcpy(dw, wfn, no * sizeof(double2), hipMemcpyHostToDevice), "hipMemcpy");
  chk(hipMemcpy(dc, cic, nc * sizeof(double2), hipMemcpyHostToDevice), "hipMemcpy");
  if (td1c_htot(c, *dtime, lfield, reinterpret_cast<const double*>(dw), reinterpret_cast<const double*>(dc),
                reinterpret_cast<double*>(dh), reinterpret_cast<double*>(dhc), nullptr)) {
    std::fprintf(stderr, "hprod_htot_: %s\n", g_error.c_str());
    std::abort();
  }
  chk(hipMemcpy(hwfn, dh, no * sizeof(double2), hipMemcpyDeviceToHost), "hipMemcpy");
  chk(hipMemcpy(hcic, dhc, nc * sizeof(double2), hipMemcpyDeviceToHost), "hipMemcpy");
}

// hprod_htot_static: hprod_htot_ without any device allocation.  The device
// copies of the host arguments live in `scratch`, a file-scope device array
// of kStaticStageElems complex values that every GPU the library runs on holds
// until the process ends.  Host code never uses the array's name as an
// address: static_stage_ptr() looks the device address up once per device
// with hipGetSymbolAddress (that device current) and keeps it.
constexpr size_t kStaticStageElems = (size_t)4 << 20;  // 64 MiB of complex128
static __device__ __align__(256) double2 scratch[kStaticStageElems];

static int static_stage_ptr(double2** out) {
  static std::mutex mu;
  static std::map<int, double2*> ptrs;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it == ptrs.end()) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(-3, "hipGetSymbolAddress returned a null address for the staging array");
    it = ptrs.emplace(dev, static_cast<double2*>(p)).first;
  }
  *out = it->second;
  return 0;
}

int hprod_htot_static(const double* dtime, const double* lfield, const double* wfn, const double* cic,
                      double* hwfn, double* hcic) {
  td1c_ctx* c = g_bound;
  if (!c) return fail(-1, "hprod_htot_static: no bound td1c context");
  if (!dtime || !lfield || !wfn || !hwfn) return fail(-1, "hprod_htot_static: null argument");
  const size_t no = (size_t)c->nbas * c->S.nfun, nc = (size_t)c->S.lcic;
  if (nc > 0 && (!cic || !hcic)) return fail(-1, "hprod_htot_static: null CI argument");
  if (2 * (no + nc) > kStaticStageElems)
    return fail(-2, "hprod_htot_static: the wave function (" + std::to_string(no + nc) +
                        " complex values) exceeds the static staging limit of " +
                        std::to_string(kStaticStageElems / 2) + " (input and output share a 64 MiB device array)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != c->device) return fail(-1, "hprod_htot_static: the bound context belongs to another device");
  double2* dw = nullptr;
  int rc = static_stage_ptr(&dw);
  if (rc) return rc;
  double2* dh = dw + no;
  double2* dc = dh + no;
  double2* dhc = dc + nc;
  HIP_TRY(hipMemcpy(dw, wfn, no * sizeof(double2), hipMemcpyHostToDevice));
  if (nc) HIP_TRY(hipMemcpy(dc, cic, nc * sizeof(double2), hipMemcpyHostToDevice));
  rc = td1c_htot(c, *dtime, lfield, reinterpret_cast<const double*>(dw), reinterpret_cast<const double*>(dc),
                 reinterpret_cast<double*>(dh), reinterpret_cast<double*>(dhc), nullptr);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(hwfn, dh, no * sizeof(double2), hipMemcpyDeviceToHost));
  if (nc) HIP_TRY(hipMemcpy(hcic, dhc, nc * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
