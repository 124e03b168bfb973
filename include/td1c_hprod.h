/*
 * td1c_hprod.h -- C-ABI of the MI355X (gfx950) orbital Hamiltonian product
 * ("hprod") and its explicit Runge-Kutta step.
 *
 * The library replaces, for one process-wide problem, the Fortran entry
 * points that td1c's C++ drivers call through include/wrapper.hpp:
 *
 *   hprod_htot_        include/wrapper.hpp:159, called by clhprod::htot
 *                      (src/hprod/clhprod.cpp:62-82)
 *   hprod_set_wfn0_    include/wrapper.hpp:400, called by clhprod::set_wfn0
 *                      (src/hprod/clhprod.cpp:50-55)
 *   bas_rad_bind_, bas_sph_bind_, bas_bas_bind_, control_bind_, ormas_bind_,
 *   hprod_bind_, hprod_init_ (src/bas/clbas.cpp:60-80,135-141;
 *                      src/control/clcontrol.cpp:185-189;
 *                      src/hprod/clhprod.cpp:44-47)
 *                      -> folded into td1c_ctx_create, which takes the same
 *                      grid / angular / orbital / control parameters by value
 *                      and builds every operator table itself.
 *   clcrk::prop        src/prop/clcrk.cpp:55-87 (rk_order 1..4)
 *                      -> td1c_rk_prop
 *   clwfn::mask        src/wfn/clwfn.cpp:337-357, src/wfn/wfn_mask.F90:2-35
 *                      -> td1c_mask
 *   clfield::get_value src/field/clfield.cpp:121-140 -> td1c_field_value
 *
 * Data layout is td1c's: an orbital vector is complex128
 * wfn[(ifun*(lmax1+1) + l)*(nrad-1) + irad], radial index fastest
 * (src/hprod/hprod_htot_dtorb.F90:15-20), and a full wave function is the
 * nfun*nbas orbital coefficients followed by the lcic CI coefficients
 * (src/wfn/clwfn.cpp:115-121).  Complex numbers are two doubles (re, im),
 * the layout of Fortran complex(c_double_complex) and std::complex<double>.
 *
 * Errors: every int-returning function returns 0 on success and a negative
 * code on failure; td1c_last_error() gives the message.  (The reference has
 * no return codes: it calls abort() or Fortran STOP.)  hprod_htot_ keeps the
 * reference's void signature and aborts with a message on failure.
 */
#ifndef TD1C_HPROD_H
#define TD1C_HPROD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Problem description: the values the reference reads from its input deck
 * (clfedvr::read_x, clsph::gen, clbas::read_info, clormas, clcontrol::gen)
 * and binds into its Fortran modules. */
typedef struct td1c_problem {
  /* radial FEDVR grid: nfe elements of ndvr+1 Gauss-Lobatto points each;
   * edges[0..nfe] are the element boundaries (src/fedvr/clfedvr.cpp:253-311) */
  int nfe;
  int ndvr;
  const double* edges;
  /* angular expansion (src/sph/clsph.cpp:97-134); nlat <= 0 means lmax2+1,
   * mmax1 < 0 means max |mval|, mmax2 < 0 means 2*mmax1 */
  int lmax1, lmax2, mmax1, mmax2, nlat;
  /* nuclear charge (src/bas/clbas.cpp:35) */
  double znuc;
  /* orbitals, ordered frozen core type -2, frozen core type -1, dynamical
   * core, active (src/ormas/ormas_info.F90:18-44); nfun is the sum */
  int nfcore2, nfcore1, ndcore, nact;
  const int* mval;          /* [nfun] magnetic quantum number per orbital */
  const int* lval;          /* [nfun] l of each orbital (frozen-core density) */
  int nel_alpha, nel_beta;  /* neltot(1:2) */
  int mtot;                 /* total M of the determinants, cores included
                               (clbas mtot; ormas_madapt_str adds the
                               cores' M to each spin string) */
  int nradfc;               /* radial extent of frozen-core work, <=0: nrad-1 */
  /* control flags (src/control/clcontrol.cpp:67-191) */
  int icomp;                /* 1: real time (tfac = -i dt), 0: imaginary time */
  int igauge;               /* 0: length1, 1: velocity1 */
  int isplit;               /* split_type: 0 full hprod; 1 the hprod's orbital
                               derivative without h0 + h1 (hprod_htot_dtorb.F90:76-83),
                               which clh1itr's half steps propagate (td1c_h1itr_*,
                               td1c_propagate); iprojfc = 2 with frozen cores */
  int ioorot;               /* oorot_type: only 0 is provided */
  int iprojfc;              /* 1 or 2 (default 2) */
  int jfc_implicit;         /* default 1 */
  int xfc_implicit;         /* default 1 */
  double throcc1, throcc2;  /* regularisation of D^-1, (2-D)^-1 (1e-15) */
  /* absorbing mask (src/fedvr/clfedvr.cpp:361-398,795-810): rmask <= 0
   * means 0.8 * rmax, mask_cos_order <= 0 means 4 */
  double rmask;
  double mask_cos_order;
  /* arithmetic of the mean field: 0 = complex128 throughout; 1 = the
   * theta-grid orbitals, pair densities and pair potentials in complex64
   * (their Legendre transforms and mean-field products in FP32, the radial
   * Poisson solve in FP64 on complex64 input/output), with the state, the
   * one-body part, the Gram products, X matrix, projector, 1-RDM inverse, CI
   * and RK accumulation in FP64 (no counterpart in the reference) */
  int precision;
  /* clbas lconst / lconst_core (clwfn::ladapt, src/wfn/clwfn.cpp:359-366):
   * 0 off (default); 1 every orbital keeps only its l = lval component
   * (wfn_ladapt, src/wfn/wfn_ladapt.F90:2-39; its CI part returns early);
   * 2 the cores only (wfn_ladapt_core, :100-125).  Applied to the orbital
   * derivative of every hprod, as clhprod::htot does to hWfn. */
  int lconst;
} td1c_problem;

/* Sizes of a created context. */
typedef struct td1c_sizes {
  int nrad;      /* FEDVR points including r=0 and r=R; interior = nrad-1 */
  int nbas;      /* (nrad-1)*(lmax1+1) */
  int nfun;
  int lcic;      /* CI length (1 when nact == 0) */
  int nlat;
  int lmax1, lmax2;
  int npair;     /* number of orbital pair densities evaluated per hprod */
  long long wfn_size;  /* nfun*nbas + lcic complex elements */
} td1c_sizes;

typedef struct td1c_ctx td1c_ctx;

/* Laser field, one pulse (src/field/clfield1.cpp:14-116):
 * env_type 0 = zero, 1 = sin2, 2 = flat; gauge 0 = length1, 1 = velocity1.
 * fint in W/cm^2, wlen in nm, cep in degrees, numcyc optical cycles. */
typedef struct td1c_field {
  int env_type;
  int gauge;
  double fint;
  double wlen;
  double cep;
  double numcyc;
  double delay;   /* pulse delay in attoseconds (converted as 1 a.u. = 24.188843265 as) */
} td1c_field;

/* Builds the FEDVR grid, angular quadrature, every operator table and the
 * device work space for one problem, on the current HIP device.
 * Replaces clbas::gen + the *_bind_ calls + hprod_init_. */
int td1c_ctx_create(const td1c_problem* prob, td1c_ctx** out);

/* Frees the context (hprod_final_). */
int td1c_ctx_destroy(td1c_ctx* ctx);

/* Sizes of a context. */
int td1c_get_sizes(const td1c_ctx* ctx, td1c_sizes* out);

/* Frozen-core set-up from the initial wave function (host pointers):
 * orb0, rhofc, v2jfc (+ v2xfc when given), tmat += v2jfc.
 * hprod_set_wfn0_ (src/hprod/hprod_set_wfn0.F90:2-130,132-219); the
 * KLI/LDA local exchange of type -2 cores is not computed here: pass it in
 * v2xfc (nrad-1 complex values, or NULL for zero), as v2xfc_type = "read"
 * does in the reference. */
int td1c_set_wfn0(td1c_ctx* ctx, const double* wfn, const double* cic,
                  const double* v2xfc);

/* One Hamiltonian product on device buffers, enqueued on `stream`
 * (a hipStream_t, NULL for the default stream):
 *   hwfn = dt * d(psi)/dt (orbitals), hcic = dt * dC/dt (CI).
 * hprod_htot (src/hprod/hprod_htot.F90:2-42) plus the madapt/ladapt
 * post-ops of clhprod::htot (madapt is a no-op; ladapt follows
 * td1c_problem.lconst). */
int td1c_htot(td1c_ctx* ctx, double dtime, const double* lfield,
              const double* d_wfn, const double* d_cic,
              double* d_hwfn, double* d_hcic, void* stream);

/* hprod_htotx_ (include/wrapper.hpp:161; src/hprod/hprod_htot.F90:44-208,
 * called by clhprod::htotx, src/hprod/clhprod.cpp:84-88): the hprod without
 * the orbital derivative (dtorb): RDMs, integrals, X matrix, energies and the
 * CI derivative stay in the context for td1c_get_* below.  On `stream`. */
int td1c_htotx(td1c_ctx* ctx, double dtime, const double* lfield,
               const double* d_wfn, const double* d_cic, void* stream);

/* The module arrays of the last td1c_htot / td1c_htotx, column-major
 * complex128 in host memory; each synchronises the device.
 * td1c_get_den1 / td1c_get_int1e: nact x nact (hprod_getden1_,
 * hprod_getint1_, include/wrapper.hpp:150-151; src/hprod/hprod_getvar.F90:2-24);
 * td1c_get_int2e / td1c_get_den2: nact^4 (hprod_getint2_, :152; :26-36, and
 * den2 of src/module/mod_hprod.F90); td1c_get_xmat: nfun x nfun, X(j,i)
 * multiplying orbital j into the derivative of orbital i, times dt. */
int td1c_get_den1(td1c_ctx* ctx, double* out);
int td1c_get_den2(td1c_ctx* ctx, double* out);
int td1c_get_int1e(td1c_ctx* ctx, double* out);
int td1c_get_int2e(td1c_ctx* ctx, double* out);
int td1c_get_xmat(td1c_ctx* ctx, double* out);

/* Energies of the last td1c_htot (ene_fcore, ene_dcore, ene_core, ene_act,
 * ene_tot as bound by hprod_bind_); synchronises the stream. */
int td1c_get_energies(td1c_ctx* ctx, double ene[5]);

/* Length-form dipole <z> of a device wave function, written to d_out (one
 * double in device memory) on `stream`: hprod_dipole's dip(1)
 * (src/hprod/hprod_dipole.F90:2-146), i.e. hprod_trace1 of z with the 1-RDM,
 * cores counted twice, frozen cores with the velocity-gauge phase of lfield. */
int td1c_dipole(td1c_ctx* ctx, const double* lfield, const double* d_wfn,
                const double* d_cic, double* d_out, void* stream);

/* Length, velocity and acceleration forms of the dipole, written to d_out
 * (three doubles in device memory) on `stream`: hprod_dipole's Ehrenfest
 * terms (src/hprod/hprod_dipole.F90:51-72) d_out = {dip(1), vel(1), acc(1)},
 * the 1-RDM traces (hprod_trace1) of the property vectors of
 * hprod_zprod_all, hprod_pzprod_all and hprod_azprod_all at zfac = 1 (bare
 * nucleus: a_z = -Z cos(theta) / r^2), plus A_z N in the velocity gauge for
 * vel and -E_z N for acc (N = all electrons).  Without frozen cores these
 * are hprod_dipole's totals dip(3), vel(3), acc(3).  With frozen cores the
 * reference adds correction terms vel(2), acc(2) built on the generalised
 * Fock vector of hprod_gfock (src/hprod/hprod_dipole.F90:82-125); they are
 * in td1c_dipole_all. */
int td1c_dipole3(td1c_ctx* ctx, const double* lfield, const double* d_wfn,
                 const double* d_cic, double* d_out, void* stream);

/* hprod_dipole in full (src/hprod/hprod_dipole.F90:2-146): d_out (nine
 * doubles in device memory) = {dip(1), dip(2), dip(3), vel(1), vel(2),
 * vel(3), acc(1), acc(2), acc(3)}, the reference's Phys.dip, vel, acc.
 * (1) are td1c_dipole3's traces; with frozen cores (2) are the corrections
 * of :74-125 on the generalised Fock vector f of hprod_gfock (the one-body
 * part, the mean field including its action on the frozen cores, plus E_z z
 * in the velocity gauge): vel(2) = -4 Im sum_fc <z i|f i> - 2 Im tr(D z,
 * P_fc f), acc(2) the same with p_z, P_fc the frozen-core projector of
 * hprod_projfc; dip(2) = 0; (3) = (1) + (2).  Costs one hprod evaluation
 * (print cadence).  Enqueued on `stream`. */
int td1c_dipole_all(td1c_ctx* ctx, const double* lfield, const double* d_wfn, const double* d_cic, double* d_out,
                    void* stream);

/* Ionisation probabilities of a device wave function: ipx[k] (host memory,
 * max_ipx + 1 doubles) = probability that exactly k electrons are outside
 * rad_ipx, k = 0..max_ipx.  Replaces hprod_ipx_ (include/wrapper.hpp:196;
 * src/hprod/hprod_ipx.F90:2-19, called by clhprod::ionpx and clhprod::normx)
 * with ormas_ipx (src/ormas/ormas_ipx.F90:2-35 -> ormas_ipx_core /
 * ormas_ipx_ras): the orbitals of hprod_orbin at lfield, their overlap over
 * the radial points up to and including the first r > rad_ipx
 * (hprod_mkovlp), and the CI vector.  0 <= max_ipx <= min(8, electrons), as
 * the reference requires (it stops otherwise; here the call returns -1).
 * Synchronises `stream`. */
int td1c_ipx(td1c_ctx* ctx, int max_ipx, double rad_ipx, const double* lfield,
             const double* d_wfn, const double* d_cic, double* ipx, void* stream);

/* hprod_ipd_ (include/wrapper.hpp:195; src/hprod/hprod_ipx.F90:21-39 ->
 * ormas_ipd, src/ormas/ormas_ipd.F90:2-123, called by clhprod::ionpd):
 * ipd[k] (host memory, max_ipd + 1 doubles, max_ipd <= min(2, electrons)) =
 * the estimate p0 = 1 - S1 + S2, p1 = S1 - 2 S2, p2 = S2 of the probability
 * that k electrons are outside rad_ipx, S1 and S2 the first two factorial
 * moments of the count (the reference forms them from the 1- and 2-RDM).
 * Synchronises `stream`. */
int td1c_ipd(td1c_ctx* ctx, int max_ipd, double rad_ipx, const double* lfield,
             const double* d_wfn, const double* d_cic, double* ipd, void* stream);

/* clwfn::orth with orth_type "schmidt" (src/wfn/clwfn.cpp:313-321;
 * wfn_orth_cic and wfn_orth_orb, src/wfn/wfn_orth.F90:2-176) on a device wave
 * function: the CI vector normalised, each dynamical orbital in order made
 * orthogonal to the earlier orbitals of the same m (modified Gram-Schmidt)
 * and normalised.  Enqueued on `stream`. */
int td1c_orth(td1c_ctx* ctx, double* d_wfn, void* stream);

/* Imaginary-time relaxation, init_split2 with split_type 0 and the clcrk
 * propagator (src/main/init.cpp:178-240): each cycle evaluates the energy
 * (hprod_energy, as td1c_htot's ene_tot) at lfield, stops when it changed by
 * less than tolene, else takes one rk_order step of dstep (no mask) and
 * calls td1c_orth.  The context must have icomp = 0.  *ene gets the last
 * energy, *ncyc the last cycle index.  Synchronises `stream`. */
int td1c_init_split2(td1c_ctx* ctx, int rk_order, double dstep, double tolene,
                     int maxcyc, const double* lfield, double* d_wfn,
                     double* ene, int* ncyc, void* stream);

/* Wave-function files in the reference's text format, host memory only (no
 * GPU call).  td1c_wfn_write replaces clwfn::write
 * (src/wfn/clwfn.cpp:268-302): the .orb file holds nrad, lmax1, mmax1, nfun
 * ("%10d" lines) then every orbital coefficient as "%25.15le%25.15le" in
 * [orbital][l][r] order (nrad-1 radial points); the .cic file holds lcic then
 * the CI coefficients.  wfn: the state as complex128 (orbitals then CI). */
int td1c_wfn_write(const char* orb_path, const char* cic_path, int nrad, int lmax1,
                   int mmax1, int nfun, int lcic, const double* wfn);

/* clwfn::read (src/wfn/clwfn.cpp:156-205): wfn (sized for nrad, lmax1, nfun,
 * lcic) is cleared and filled from the files; orbitals beyond the current
 * sizes are dropped, missing ones stay zero.  The CI vector is read when the
 * file's lcic matches; otherwise determinant cic0 is set to 1, as ormas_cic0
 * does (src/ormas/ormas_cic0.F90:24; td1c_ci_layout gives cic0), and the call
 * returns 1.  cic0 < 0 (no active orbitals, where ormas_cic0 returns at
 * once) leaves the CI vector zero. */
int td1c_wfn_read(const char* orb_path, const char* cic_path, int nrad, int lmax1,
                  int nfun, int lcic, int cic0, double* wfn);

/* cletdrb (src/prop/cletdrb.cpp), the reference's default prop_type:
 * exponential Rosenbrock RK4 (prop4_orb, :321-434) with the [2/3] Pade forms
 * of phi1, phi2, phi3 (gen_phi1 / gen_phid1 / gen_phi2 / gen_phi3,
 * :636-1185) applied to the orbitals as rational functions of h1 = T + V(t)
 * (clh1rat, src/prop/clh1rat.cpp; h1rat_gen / h1rat_prod,
 * src/prop/h1rat_gen.F90, h1rat_prod.F90): numerators by applications of h1,
 * each denominator factor cd0 + cd1 (T + V) by the element Schur solve of
 * cd0 + cd1 T and the Neumann series in V (at most maxcyc terms, RMS
 * threshold thresh; h1rat_maxcyc 20 and h1rat_thresh 1e-10 by default,
 * src/control/clcontrol.cpp:123-124).  The settings are cletdrb::gen's
 * defaults (:25-92): etd_rk_order 4, etd_dim_numer* 2, etd_dim_denom* 3,
 * etd_time_of_h 0, etd_cisplit -1; rosenbrock is etd_rosenbrock (default
 * 1).  Needs split_type = 1 and lconst = 0; with frozen cores call after
 * td1c_set_wfn0 (the factors include the frozen-core Hartree potential).
 * td1c_etdrb_gen builds the tables for step dtime (cletdrb::gen). */
int td1c_etdrb_gen(td1c_ctx* ctx, double dtime, int rosenbrock, int maxcyc, double thresh);

/* One cletdrb step of the device wave function (orbitals then CI) with the
 * field arrays given on the host, lfields[6][9]: lfield(t), lfield(t + dt/2),
 * lfield(t + dt), d lfield/dt at t (td1c_field_der), lfield(t + dt/2) -
 * lfield(t), lfield(t + dt) - lfield(t); then the mask when apply_mask != 0,
 * as tdse_prop12 does after every step (src/main/tdse.cpp:268-309). */
int td1c_etdrb_prop(td1c_ctx* ctx, const double* lfields, double* d_wfn, int apply_mask, void* stream);

/* nstep cletdrb steps of tdse_prop12 from t0 with the field evaluated on the
 * device; dtime must be td1c_etdrb_gen's. */
int td1c_etdrb_propagate(td1c_ctx* ctx, const td1c_field* field, double t0, double dtime, int nstep,
                         double* d_wfn, int apply_mask, void* stream);

/* cletdrk4 (src/prop/cletdrk4.cpp, prop_type "etdrk4"): the Cox-Matthews
 * exponential RK4 (prop, :69-114) with the [2/2] Pade forms of its defaults
 * (etd_dim_numer = etd_dim_denom = 2, gen :24-60) for exp(dt h1),
 * exp(dt h1 / 2), phi1 (gfun) and the three weight functions f0fun, f1fun,
 * f2fun (gen_exp ... gen_f2fun, :134-933), all taking the field at t + dt/2;
 * the CI part is scaled by the functions' dt -> 0 limits (exp1rk ...
 * f2funrk).  The rational functions are solved as td1c_etdrb_gen's; the same
 * restrictions apply.  A context holds one scheme's tables at a time
 * (td1c_etdrb_gen or td1c_etdrk4_gen, the last call wins); lfields as for
 * td1c_etdrb_prop (this scheme reads lfield(t), lfield(t + dt/2) and
 * lfield(t + dt)). */
int td1c_etdrk4_gen(td1c_ctx* ctx, double dtime, int maxcyc, double thresh);
int td1c_etdrk4_prop(td1c_ctx* ctx, const double* lfields, double* d_wfn, int apply_mask, void* stream);
int td1c_etdrk4_propagate(td1c_ctx* ctx, const td1c_field* field, double t0, double dtime, int nstep,
                          double* d_wfn, int apply_mask, void* stream);

/* cllawrk (src/prop/cllawrk.cpp, prop_type "lawrk"): Lawson RK4 (prop4,
 * :188-280) with exp(-i h1 dt/2) in the (1,1) Pade form of cldpade
 * (lawrk_npade 11, h1itr_midpt true: expwfn11, :296-324): out = 2 (1 + i T
 * dt/4)^-1 in - in on the orbitals, the inverse with the Neumann series in V
 * of dpade_prod (src/prop/dpade_prod.F90; at most maxcyc terms, RMS
 * threshold thresh: h1itr_maxcyc, h1itr_thresh), the field at t + dt/2.
 * The factor is the one td1c_h1itr_gen builds for this dt.  split_type = 1
 * and lconst = 0; lfields as for td1c_etdrb_prop (lfield(t), lfield(t +
 * dt/2), lfield(t + dt) are read).  A context holds one exponential scheme at
 * a time. */
int td1c_lawrk_gen(td1c_ctx* ctx, double dtime, int maxcyc, double thresh);
int td1c_lawrk_prop(td1c_ctx* ctx, const double* lfields, double* d_wfn, int apply_mask, void* stream);
int td1c_lawrk_propagate(td1c_ctx* ctx, const td1c_field* field, double t0, double dtime, int nstep,
                         double* d_wfn, int apply_mask, void* stream);

/* clexpab (src/prop/clexpab.cpp, prop_type "expab"): the exponential
 * Adams-Bashforth method of prop_general (:370-496) with gen's defaults
 * (:25-110): order K = 4, the [1/2] form of phi4 consistent with the (1,2)
 * Pade of exp (gen_phi4, :1148-1354), no corrector, h1type -1.  One split
 * hprod per step; the history of the last K states and derivatives (filled
 * with the start state on the first step after td1c_expab_gen) lives in the
 * context, so consecutive calls continue one propagation.
 * td1c_expab_prop: one step at time t with lfields[4][9] = lfield(t), then
 * lfield at the history times timeP[1..3] (t_(n-k) for the n-th step since
 * gen, t_0 while n < k).  td1c_expab_propagate evaluates them itself. */
int td1c_expab_gen(td1c_ctx* ctx, double dtime, int maxcyc, double thresh);
int td1c_expab_prop(td1c_ctx* ctx, double t, const double* lfields, double* d_wfn, int apply_mask,
                    void* stream);
int td1c_expab_propagate(td1c_ctx* ctx, const td1c_field* field, double t0, double dtime, int nstep,
                         double* d_wfn, int apply_mask, void* stream);

/* hprod_chkconv (src/hprod/hprod_chkconv.F90:2-63), the convergence test of
 * init_prop12: out[0] the energy of an hprod at lfield, out[1] |hcic| /
 * lcic, and from the generalised Fock vectors (h0 + h1 + g + v2) phi_i of
 * the dynamical orbitals (hprod_gfock), projected on the complement of all
 * orbitals (hprod_projall, inner false) and overlapped over the whole box
 * (hprod_mkovlp): out[2] = sqrt(sum_i <g_i|g_i>) / nfun, out[3] =
 * sqrt(2 sum_dcore <g_i|g_i> + sum_act <g_a|g_b> den1(b, a)).  Synchronises
 * the stream.  Needs nradfc = nrad-1. */
int td1c_chkconv(td1c_ctx* ctx, const double* lfield, const double* d_wfn, double out[4], void* stream);

/* init_prop12 (src/main/init.cpp:242-310) with cletdrb, the default
 * prop_type: until |E - E_prev| < tolene and out[3] of td1c_chkconv <
 * tolwfn (init_tolene 1e-12, init_tolwfn 1e-6 by default), one cletdrb step
 * of dstep at the fixed field lfield (icomp = 0) then clwfn::orth.  The
 * context must have icomp = 0 and split_type = 1; the series settings are
 * td1c_h1itr_gen's (20 terms, 1e-10 by default).  *ene gets the last energy,
 * *ncyc the last cycle index. */
int td1c_init_prop12(td1c_ctx* ctx, double dstep, double tolene, double tolwfn, int maxcyc,
                     const double* lfield, double* d_wfn, double* ene, int* ncyc, void* stream);

/* Grid projection of a state between two problems, host memory only (no GPU
 * call): clbas::proj (src/bas/clbas.cpp:280-332) with the CI copy of
 * guess_type = "read_proj" (src/main/guess.cpp:37-47).  Each interior point r
 * of `to`'s grid gets sqrt(w(r)) * sum_m c_m L_m(r) over the stored points of
 * the `from` element holding r (the first whose right edge exceeds r + 1e-10,
 * clfedvr::get_ife, src/fedvr/clfedvr.cpp:628-638), L_m the element's
 * Lagrange polynomials (cldvr::get_val, src/dvr/cldvr.cpp:223-235) and c_m the
 * source coefficients as stored.  Orbitals and l up to the smaller of the two
 * problems' counts are mapped (the reference loops to the larger and then
 * reads across orbital blocks); points beyond `from`'s grid stay zero (the
 * reference indexes element -1 there).  The first min(lcic) CI coefficients
 * are copied, the rest are zero.  wfn_from / wfn_to: complex128 states
 * (orbitals then CI) of the two problems.  The reference then calls
 * clwfn::orth: td1c_orth on the device copy. */
int td1c_wfn_proj(const td1c_problem* from, const double* wfn_from,
                  const td1c_problem* to, double* wfn_to);

/* Layout of the CI vector of a problem, computed on the host (no device, no
 * context): the reference's ORMAS addressing for one subspace.  Strings of
 * one spin are numbered by arc weights (ormas_arcwgt,
 * src/ormas/ormas_arcwgt.F90:52-104; decoded by ormas_occvec,
 * src/ormas/ormas_occvec.F90:2-29) and reordered by their total M, cores
 * included (ormas_madapt_str, src/ormas/ormas_madapt.F90:170-246); the CI
 * vector holds the determinants with M_alpha + M_beta = mtot in the order
 * ntot_alph_beta(beta) + alpha (:381-435), beta major.  Outputs (each may be
 * NULL): *nstr_alpha, *nstr_beta, *lcic, *cic0 (the determinant ormas_cic0
 * sets, src/ormas/ormas_cic0.F90:24; -1 when nact == 0; >= lcic when the
 * first alpha and beta strings together miss M = mtot, where ormas_cic0
 * would write past the CI vector and td1c_wfn_read reports an error instead); str_alpha /
 * str_beta (nstr_* entries): occupation bit patterns of the ordered strings,
 * bit k = active orbital k; det_alpha / det_beta (lcic entries): the string
 * indices of each determinant.  Call once with NULL arrays for the sizes. */
int td1c_ci_layout(const td1c_problem* prob, int* nstr_alpha, int* nstr_beta, int* lcic,
                   int* cic0, unsigned int* str_alpha, unsigned int* str_beta,
                   int* det_alpha, int* det_beta);

/* One explicit Runge-Kutta step of the whole wave function (orbitals then
 * CI) held on device: clcrk::prop with rk_order 1..4 ("standard").
 * lfield_stages holds the 9 field values at t + c_s*dt for each stage
 * (host memory, nstage*9 doubles).  apply_mask != 0 applies clwfn::mask
 * afterwards, as tdse_split2 does after every step (src/main/tdse.cpp:225-266). */
int td1c_rk_prop(td1c_ctx* ctx, int rk_order, double dtime,
                 const double* lfield_stages, double* d_wfn, int apply_mask,
                 void* stream);

/* clh1itr::gen (src/prop/clh1itr.cpp:29-59; h1itr_gen_cninv,
 * src/prop/h1itr_gen_cninv.F90:2-58) for a context with split_type = 1:
 * the factors of the one-body half steps of a step of dtime (half steps of
 * dtime/2: Crank-Nicolson with 1 + i T dtime/4, T = tmat including the
 * frozen-core Hartree potential), with the reference's h1itr_maxcyc
 * (default 20) and h1itr_thresh (default 1e-10).  td1c_propagate builds the
 * factors itself (with the last maxcyc / thresh given, else the defaults)
 * when dtime or the frozen core changed. */
int td1c_h1itr_gen(td1c_ctx* ctx, double dtime, int maxcyc, double thresh);

/* One one-body half step, clh1itr::prop -> h1itr_prop
 * (src/prop/h1itr_prop.F90:2-134, iprojfc = 2): the dynamical orbitals of
 * the device wave function are replaced by
 *   sum_k x_k,  x_0 = (1 + c T)^-1 (1 - c (T + V)) psi,
 *               x_k = (1 + c T)^-1 (-c V) x_(k-1),
 * c = i dtime/4, V the dipole coupling of lfield[2] in the context's gauge,
 * summed while the RMS orbital norm of the last term is >= h1itr_thresh and
 * for at most h1itr_maxcyc terms after x_0.  lfield is the field the
 * reference evaluates at t + dt/4 (first half step) or t + 3dt/4 (second).
 * The first check of the series (on x_0, norm ~1 per orbital) counts the
 * dynamical orbitals only.  Enqueued on `stream`; td1c_h1itr_gen first.
 * Under stream launches the call reads the series' flag back after the
 * second term and every two terms after (it waits for `stream` to get
 * there) and enqueues no more terms once the series has converged; under
 * graph capture every term is enqueued.  The same holds for the half steps
 * of td1c_propagate and the denominator series of the exponential
 * propagators. */
int td1c_h1itr_prop(td1c_ctx* ctx, const double* lfield, double* d_wfn, void* stream);

/* nstep steps (+ mask) from time t0 with the stage fields evaluated on the
 * device: the time loop of tdse_split2 (src/main/tdse.cpp:225-266) without
 * its printing.  split_type = 0: each step does what td1c_rk_prop does with
 * clfield::get_value's values (computed on the device from the pulse
 * parameters, no host round trip).  split_type = 1: each step is
 * td1c_h1itr_prop with the field at t + dt/4, td1c_rk_prop of the split
 * hprod, td1c_h1itr_prop at t + 3dt/4, then the mask.  With
 * td1c_set_graph(ctx, 1) the first call for a given (rk_order, dtime,
 * field, d_wfn, apply_mask) captures one step as a HIP graph and every step
 * replays it.  Enqueued on `stream`, asynchronous. */
int td1c_propagate(td1c_ctx* ctx, const td1c_field* field, int rk_order,
                   double t0, double dtime, int nstep, double* d_wfn,
                   int apply_mask, void* stream);

/* Batches of independent propagations (an intensity / wavelength / CEP scan
 * over grids and lmax): the reference runs one tdse_split2 time loop
 * (src/main/tdse.cpp:225-266) per problem, each in its own process; a batch
 * runs the same steps for many problems at once, each member's kernels
 * issued as one launch per kernel site for the whole batch.  Members are
 * contexts the caller created (and keeps; destroy the batch first) on one
 * device with ndvr = 20, dynamical cores only (no frozen cores, no active
 * orbitals), and the same nfun, precision, split_type, gauge and icomp;
 * their grids, lmax, nuclear charges and pulses may differ.  The batch
 * uses the matrix-core form of every transform (td1c_set_transform_path 2
 * gives a member's own td1c_propagate the same bits). */
typedef struct td1c_batch td1c_batch;
int td1c_batch_create(td1c_ctx* const* members, int n, td1c_batch** out);
int td1c_batch_destroy(td1c_batch* batch);

/* nstep steps of every member from t0: member m's wave function d_wfns[m]
 * (device) under fields[m], exactly as td1c_propagate(members[m],
 * &fields[m], rk_order, t0, dtime, nstep, d_wfns[m], apply_mask, stream)
 * would step it.  The first call for a given set of arguments records one
 * step of every member and uploads the launch table; with graphs on (the
 * default) one step of the whole batch is captured once and replayed per
 * step.  Enqueued on `stream`, asynchronous. */
int td1c_batch_propagate(td1c_batch* batch, const td1c_field* fields, int rk_order, double t0, double dtime,
                         int nstep, double* const* d_wfns, int apply_mask, void* stream);

/* 1 (default): one captured HIP graph per batch step; 0: stream launches. */
int td1c_batch_set_graph(td1c_batch* batch, int on);

/* Launch shapes of the batch: the element one-body kernel's elements per
 * block (3, 6, 12) and l values per block, and the register-blocked
 * transform's variant (0: 4 column tiles, 1: 2 column tiles per wavefront);
 * every choice gives the same bits. */
int td1c_batch_set_shapes(td1c_batch* batch, int elements_per_block, int l_per_block, int transform_variant);

/* Chooses the batch's launch shapes by timing: every element one-body shape
 * (3, 6, 12 elements x 2, 4, 8, 16 l values per block), then the other
 * transform variant, each as one warm and two timed batch steps of dtime
 * on the scratch wave functions d_scratch[m] (device copies the caller
 * provides; they are overwritten).  Synchronises the stream. */
int td1c_batch_autotune(td1c_batch* batch, const td1c_field* fields, int rk_order, double dtime,
                        double* const* d_scratch, void* stream);

/* JSON: members, launch sites and blocks per step, table bytes, shapes, the
 * members reading shared element tables (shared_tables: cn, poisson,
 * onebody, poisson_rows), the item groups of the grouped Poisson and
 * half-step interior passes (poisson_groups, cn_groups; 0 when not
 * grouped), the autotuner's timings. */
int td1c_batch_get_info(const td1c_batch* batch, char* buf, int buflen);

/* How td1c_propagate launches its steps: 0 (default) enqueues every kernel
 * of every step on the stream, so that the independent branches of an hprod
 * (CI density stage, one-body part, pair Gram products) run concurrently on
 * the context's side streams; 1 captures one step as a HIP graph and
 * replays it (fewer host calls, the branches serialised). */
int td1c_set_graph(td1c_ctx* ctx, int on);

/* Which form of the theta <-> l transforms the context's kernels use:
 * 0 (default) chooses per launch by grid size, 1 always the vector form,
 * 2 always the matrix-core (MFMA) form.  The results agree to rounding.
 * No counterpart in the reference (its transforms are the matrix products
 * of src/bas/bas_sph2ang1.F90, bas_sph2ang2.F90, bas_ang2sph1.F90 and
 * bas_ang2sph2.F90); provided so that either form can be run and checked
 * on any problem size. */
int td1c_set_transform_path(td1c_ctx* ctx, int path);

/* Applies the radial mask to the dynamical orbitals of a device wave
 * function (wfn_mask, src/wfn/wfn_mask.F90:2-35). */
int td1c_mask(td1c_ctx* ctx, double* d_wfn, void* stream);

/* Field value at time t: lfield[0..2] chosen gauge, [3..5] E, [6..8] A
 * (src/field/clfield1.cpp:131-154). t is measured from the pulse start. */
int td1c_field_value(const td1c_field* field, double t, double lfield[9]);

/* Time derivative of the field at t in dfield[0..2] (clfield::get_der,
 * src/field/clfield.cpp:142-154, clfield1::get_der,
 * src/field/clfield1.cpp:193-211): the analytic dE/dt in the length gauge,
 * dA/dt = -E (the 6-point difference of get_evalue_numder) in the velocity
 * gauge; zero outside the pulse. */
int td1c_field_der(const td1c_field* field, double t, double dfield[9]);

/* Optical period (a.u.) of the field, used to set dtime = period/nstep. */
double td1c_field_period(const td1c_field* field);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every kernel the context enqueues (no counterpart in the reference, which
 * only times whole runs with clock(), src/main/tdse.cpp:128-129).
 * td1c_set_timing(ctx, 1) starts collecting (and clears), 0 stops;
 * td1c_get_timing writes {"kernel": [total_ms, launches], ...} as JSON and
 * clears the collected records. */
int td1c_set_timing(td1c_ctx* ctx, int on);
int td1c_get_timing(td1c_ctx* ctx, char* buf, int buflen);
/* Restricts the timing to the kernel of that name (NULL: every kernel), so
 * that a timed run carries events around one kernel only. */
int td1c_set_timing_kernel(td1c_ctx* ctx, const char* name);

/* Launch shapes the autotuner chose for this context (no counterpart in
 * the reference): at td1c_ctx_create the element one-body kernel's
 * candidates (3, 6 or 12 elements per block x 1..16 l values per block)
 * are timed on the context's buffers and the fastest kept, cached per
 * problem shape; all candidates give the same bits.  TD1C_AUTOTUNE=0 in the
 * environment keeps the defaults.  Writes JSON with the choice and the
 * measured microseconds per launch. */
int td1c_get_shapes(const td1c_ctx* ctx, char* buf, int buflen);
/* Forces the element one-body kernel's launch shape (3, 6 or 12 elements
 * per block; 1..64 l values per block), so that tests run every candidate
 * (no counterpart in the reference). */
int td1c_set_onebody_shape(td1c_ctx* ctx, int elements_per_block, int l_per_block);

/* Binds a context as the process-wide one used by hprod_htot_ (the
 * reference allows a single clhprod per process, clhprod.cpp:32-34). */
int td1c_bind(td1c_ctx* ctx);

/* Drop-in for the Fortran entry point with the reference's exact argument
 * list (include/wrapper.hpp:159): host pointers, synchronous, operates on
 * the bound context.  wfn/hwfn hold nfun*nbas complex values, cic/hcic lcic. */
void hprod_htot_(const double* dtime, const double* lfield,
                 const double* wfn, const double* cic,
                 double* hwfn, double* hcic);

/* hprod_htot_ for callers that cannot allocate device memory (code that
 * must not call hipMalloc: tight serving loops, harnesses that forbid it).
 * No counterpart in the reference.  Same arguments and results as
 * hprod_htot_ on the bound context; the device copies of the four host
 * arrays live in one static 64 MiB device array per GPU, held until the
 * process ends, so the call allocates nothing.  Limit: 2*(nfun*nbas + lcic)
 * complex values <= 4194304 (a wave function of at most 2097152 complex
 * values); larger problems return -2.  Calls that overlap on one GPU (other
 * threads) share the array and race.  Returns 0 or a negative code
 * (td1c_last_error). */
int hprod_htot_static(const double* dtime, const double* lfield,
                      const double* wfn, const double* cic,
                      double* hwfn, double* hcic);

/* Message of the last error on this thread. */
const char* td1c_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TD1C_HPROD_H */
