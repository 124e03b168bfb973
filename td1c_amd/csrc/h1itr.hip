// gfx950 kernels of the split-operator one-body half step of td1c's split2
// propagation with split_type = 1 (clh1itr, src/prop/clh1itr.cpp; h1itr_prop,
// src/prop/h1itr_prop.F90:2-134):
//   x0  = (1 + c T)^-1 (1 - c (T + V)) psi            (c = i dt/4, or dt/4)
//   x_k = (1 + c T)^-1 (-c V) x_(k-1)                 while rms(x_(k-1)) >= thresh
//   psi = sum_k x_k                                   (dynamical orbitals)
// T is tmat (kinetic band + nuclear + centrifugal + frozen-core Hartree), V
// the length- or velocity-gauge coupling of l and l +- 1.  The reference
// solves with zgbtrf/zgbtrs on the whole band per l (h1itr_cninv); here the
// same operator is solved by element Schur complement: the interior blocks
// of each FEDVR element are eliminated with their precomputed inverses and
// the bridge points form a tridiagonal system solved by cyclic reduction.
// One term of the series is four launches: the right-hand side (the
// element one-body kernel in its Crank-Nicolson modes, or a band kernel for
// ndvr != 20), the bridge pass (one block per orbital row), the interior
// pass (one block per (l, element chunk), each element inverse read once for
// every orbital) and the convergence test (one block); every later launch of
// the half step returns at once after convergence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "devmath.hpp"
#include "kernels.hpp"

namespace td1c {

// Right-hand side, thread per (r, l, dynamical orbital):
// mode 0: y = psi - c (T psi + V psi)   (V with isplit = 1 only;
//         hprod_tprod_dyn + hprod_zprod_dyn / hprod_pzprod_dyn, then the
//         zaxpy of h1itr_prop.F90:57-70)
// mode 1: y = -c V x                    (h1itr_prop.F90:106-114)
// mode 3: y = zfac1 (T x + V x)         (hprod_h1tot, hprod_h1tot.F90:2-53)
__global__ void __launch_bounds__(256) k_cn_rhs(const CnArgs a, int mode) {
  if (mode == 1 && a.flag[0]) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  const int fd = blockIdx.z;
  const int f = a.nfcore + fd;
  if (r >= a.nr) return;
  const int m = a.mval[f];
  const int am = m < 0 ? -m : m;
  double2* out = a.y + ((size_t)fd * a.L1 + l) * a.nr + r;
  if (l < am) {
    *out = make_double2(0.0, 0.0);
    return;
  }
  const double2* src = mode == 0 ? a.wfn + (size_t)f * a.L1 * a.nr : a.x + (size_t)fd * a.L1 * a.nr;
  const bool tpart = mode == 0 || mode == 3;
  const double2* psi_l = src + (size_t)l * a.nr;
  const int n = a.n;
  const int jlo = max(0, r - n), jhi = min(a.nr - 1, r + n);
  const double xr = a.xr[r];
  const double* al = a.alph + (size_t)(m + a.mmax1) * a.L1;
  double2 h = make_double2(0.0, 0.0);
  if (tpart) {
    const double* band = l == 0 ? a.band_k : a.band_k0;
    for (int j = jlo; j <= jhi; ++j) {
      const int d = j - r + n;
      double2 t;
      if (d == n) {
        t = a.tdiag[(size_t)l * a.nr + r];
      } else {
        t = make_double2(band[(size_t)d * a.nr + r], 0.0);
        if (t.x == 0.0) continue;
      }
      cfma(h, psi_l[j], t);
    }
  }
  if (mode != 0 || a.isplit == 1) {
    // the field value: lfield(3) in modes 0 and 3, -c lfield(3) in mode 1
    const double2 zf = mode == 1 ? cmul(a.zfac1, make_double2(a.lf[2], 0.0)) : make_double2(a.lf[2], 0.0);
    if (a.igauge == 0) {
      if (l + 1 < a.L1) cfma(h, psi_l[a.nr + r], cscale(zf, xr * al[l]));
      if (l - 1 >= am) cfma(h, psi_l[-(ptrdiff_t)a.nr + r], cscale(zf, xr * al[l - 1]));
    } else {
      if (l + 1 < a.L1) {
        const double2* p1 = psi_l + a.nr;
        double2 dw = make_double2(0.0, 0.0);
        for (int j = jlo; j <= jhi; ++j) {
          const double pm = a.band_p[(size_t)(j - r + n) * a.nr + r];
          if (pm == 0.0) continue;
          dw.x = fma(p1[j].x, pm, dw.x);
          dw.y = fma(p1[j].y, pm, dw.y);
        }
        cfma(h, dw, cmul(zf, make_double2(0.0, -al[l])));
        cfma(h, p1[r], cmul(zf, make_double2(0.0, -(al[l] * (l + 1.0)) / xr)));
      }
      if (l - 1 >= am) {
        const double2* p0 = psi_l - a.nr;
        double2 dw = make_double2(0.0, 0.0);
        for (int j = jlo; j <= jhi; ++j) {
          const double pm = a.band_p[(size_t)(j - r + n) * a.nr + r];
          if (pm == 0.0) continue;
          dw.x = fma(p0[j].x, pm, dw.x);
          dw.y = fma(p0[j].y, pm, dw.y);
        }
        cfma(h, dw, cmul(zf, make_double2(0.0, -al[l - 1])));
        cfma(h, p0[r], cscale(cmul(zf, make_double2(0.0, -(al[l - 1] * (double)l) / xr)), -1.0));
      }
    }
  }
  if (mode == 0) {
    double2 y = psi_l[r];
    cfma(y, a.zfac1, h);
    *out = y;
  } else if (mode == 3) {
    *out = cmul(a.zfac1, h);
  } else {
    *out = h;
  }
}

// Bridge pass, one block per (dynamical orbital, l): the Schur right-hand
// sides y_B - W_r(k) . y_I(k) - W_l(k+1) . y_I(k+1) (four lanes per bridge),
// cyclic reduction of the bridge system in LDS, the bridge values into x and
// into the state, and the block's share of sum |x|^2.
template <int BS>
__device__ __forceinline__ void cn_bridge_body(const CnArgs& a, int mode, const Blk bk) {
  if (mode == 1 && a.flag[0]) return;
  const int tid = threadIdx.x;
  // a new series: no launch of the first term reads the flag
  if (mode == 0 && bk.x == 0 && tid == 0) a.flag[0] = 0;
  const int job = bk.x;  // fd * L1 + l
  const int fd = job / a.L1, l = job - fd * a.L1;
  const int f = a.nfcore + fd;
  const int nb = a.nb, n = a.n, ni = a.ni, K = a.K, nr = a.nr;
  __shared__ double2 sS[kMaxBridge];
  __shared__ double2 sY[kMaxBridge];
  __shared__ double sRed[BS / 64];
  const int m = a.mval[f];
  const int am = m < 0 ? -m : m;
  double nrm = 0.0;
  if (l >= am) {  // uniform over the block
    const size_t row = ((size_t)fd * a.L1 + l) * nr;
    const double2* y = a.y + row;
    const double2* wl = a.Wl + (size_t)l * nr;
    const double2* wr = a.Wr + (size_t)l * nr;
    // a quad of lanes per bridge; 4 nb and the stride are multiples of 4, so
    // a quad stays together in the loop
    for (int t4 = tid; t4 < 4 * nb; t4 += BS) {
      const int kb = t4 >> 2, sub = t4 & 3;
      const int bl = kb * n, br = (kb + 1) * n;  // first interior rows of elements kb and kb+1
      double2 cl = make_double2(0.0, 0.0), cr = make_double2(0.0, 0.0);
      for (int i = sub; i < ni; i += 4) {
        cfma(cr, wr[bl + i], y[bl + i]);
        cfma(cl, wl[br + i], y[br + i]);
      }
      for (int o = 1; o < 4; o <<= 1) {
        cl.x += __shfl_xor(cl.x, o, 64);
        cl.y += __shfl_xor(cl.y, o, 64);
        cr.x += __shfl_xor(cr.x, o, 64);
        cr.y += __shfl_xor(cr.y, o, 64);
      }
      if (sub == 0) sY[kb] = csub(csub(y[br - 1], cl), cr);
    }
    __syncthreads();
    double2* cur = sY;
    double2* nxt = sS;
    for (int k = 0, h = 1; k < K; ++k, h <<= 1) {
      if (tid < nb) {
        const int i = tid;
        double2 d = cur[i];
        if (i - h >= 0) cfma(d, a.pa[((size_t)l * K + k) * nb + i], cur[i - h]);
        if (i + h < nb) cfma(d, a.pg[((size_t)l * K + k) * nb + i], cur[i + h]);
        nxt[i] = d;
      }
      __syncthreads();
      double2* t = cur;
      cur = nxt;
      nxt = t;
    }
    if (tid < nb) {
      const int gi = (tid + 1) * n - 1;
      const double2 x = cmul(cur[tid], a.binv[(size_t)l * nb + tid]);
      a.x[row + gi] = x;
      double2* w = a.wfn + ((size_t)f * a.L1 + l) * nr + gi;
      *w = mode == 0 ? x : cadd(*w, x);
      nrm = x.x * x.x + x.y * x.y;
    }
  }
  for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o, 64);
  if ((tid & 63) == 0) sRed[tid >> 6] = nrm;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) s += sRed[w];
    a.part[job] = s;
  }
}

template <int BS>
__global__ void __launch_bounds__(BS) k_cn_bridge(const CnArgs a, int mode) {
  cn_bridge_body<BS>(a, mode, hw_blk());
}

// Interior pass, one block per (l, chunk of epb elements), thread per
// interior row, every dynamical orbital in groups of kCnGroup: x_I = G_e y_I
// - W_l x_left - W_r x_right with the row of G_e loaded once into registers
// for all orbitals, into x and the state, and the block's share of sum
// |x|^2 into the partial sums that k_cn_check adds (the test at the top of
// the next cycle of h1itr_prop.F90:75-99).
template <int NI>
__device__ __forceinline__ void cn_interior_body(const CnArgs& a, int mode, int epb, int nchunk, const Blk bk) {
  if (mode == 1 && a.flag[0]) return;
  constexpr int PG = kCnGroup;
  constexpr int BS = kCnInteriorThreads;
  const int ni = NI > 0 ? NI : a.ni;
  const int nb = a.nb, nr = a.nr, n = a.n, L1 = a.L1;
  const int nd = a.nfun - a.nfcore;
  const int job = xcd_remap(bk.x, bk.nx);
  const int l = job / nchunk, ch = job - l * nchunk;
  const int e0 = ch * epb;
  const int t = threadIdx.x;
  const int el = t / ni, i = t - el * ni;
  const int e = e0 + el;
  const bool row_ok = el < epb && e < a.nfe;
  const int r = e * n + i;
  __shared__ double2 sb[PG][BS];
  __shared__ double2 sx[PG][34];  // bridges e0-1 .. e0+epb-1 (epb <= 32)
  __shared__ int sval[PG];
  __shared__ double sRed[BS / 64];
  double2 gk[NI > 0 ? NI : 1];
  double2 wl = make_double2(0.0, 0.0), wr = make_double2(0.0, 0.0);
  const double2* ge = a.G + (((size_t)l * a.gstride) + (row_ok ? e : 0)) * ni * ni + i;
  if (row_ok) {
    if (NI > 0) {
#pragma unroll
      for (int k = 0; k < NI; ++k) gk[k] = ge[k * NI];
    }
    wl = a.Wl[(size_t)l * nr + r];
    wr = a.Wr[(size_t)l * nr + r];
  }
  double nrm = 0.0;
  for (int fd0 = 0; fd0 < nd; fd0 += PG) {
    __syncthreads();  // the previous group's staged rows are no longer read
    if (t < PG) {
      const int fd = fd0 + t;
      int ok = 0;
      if (fd < nd) {
        const int m = a.mval[a.nfcore + fd];
        ok = l >= (m < 0 ? -m : m);
      }
      sval[t] = ok;
    }
#pragma unroll
    for (int j = 0; j < PG; ++j) {
      const int fd = fd0 + j;
      if (row_ok && fd < nd) sb[j][t] = a.y[((size_t)fd * L1 + l) * nr + r];
    }
    for (int q = t; q < PG * (epb + 1); q += BS) {
      const int j = q / (epb + 1), k = q - j * (epb + 1);
      const int fd = fd0 + j, bidx = e0 - 1 + k;
      sx[j][k] = (fd < nd && bidx >= 0 && bidx < nb) ? a.x[((size_t)fd * L1 + l) * nr + (size_t)(bidx + 1) * n - 1]
                                                     : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (row_ok) {
      const int eb = el * ni;
#pragma unroll 1
      for (int j = 0; j < PG; ++j) {
        const int fd = fd0 + j;
        if (fd >= nd) break;
        if (!sval[j]) continue;
        double2 x = make_double2(0.0, 0.0);
        if (NI > 0) {
#pragma unroll
          for (int k = 0; k < NI; ++k) cfma(x, gk[k], sb[j][eb + k]);
        } else {
          for (int k = 0; k < ni; ++k) cfma(x, ge[k * ni], sb[j][eb + k]);
        }
        if (e >= 1) x = csub(x, cmul(wl, sx[j][el]));
        if (e <= a.nfe - 2) x = csub(x, cmul(wr, sx[j][el + 1]));
        a.x[((size_t)fd * L1 + l) * nr + r] = x;
        double2* w = a.wfn + ((size_t)(a.nfcore + fd) * L1 + l) * nr + r;
        *w = mode == 0 ? x : cadd(*w, x);
        nrm += x.x * x.x + x.y * x.y;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o, 64);
  if ((t & 63) == 0) sRed[t >> 6] = nrm;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) s += sRed[w];
    a.part[a.npart1 + bk.x] = s;
  }
}

// Convergence test of the term just formed (the test at the top of the next
// cycle of h1itr_prop.F90:75-99): the partial norms of the bridge pass
// (npart1 of them) and of the nblk interior blocks summed in a fixed order,
// flag[0] set when the RMS orbital norm is below thresh.  A launch of its own
// after the interior pass, so that the stream order makes the partial sums
// visible (a last-block ticket would need a device-scope release fence in
// every block of the interior pass).
__device__ __forceinline__ void cn_check_body(const CnArgs& a, int mode, int nblk) {
  if (mode == 1 && a.flag[0]) return;
  constexpr int BS = kCnInteriorThreads;
  __shared__ double sRed[BS / 64];
  const int t = threadIdx.x;
  const int ntot = a.npart1 + nblk;
  double s = 0.0;
  for (int q = t; q < ntot; q += BS) s += a.part[q];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((t & 63) == 0) sRed[t >> 6] = s;
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) tot += sRed[w];
    const double rms = sqrt(tot / a.nfun);
    a.flag[0] = rms < a.thresh ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kCnInteriorThreads) k_cn_check(const CnArgs a, int mode, int nblk) {
  cn_check_body(a, mode, nblk);
}

template <int NI>
__global__ void __launch_bounds__(kCnInteriorThreads) k_cn_interior(const CnArgs a, int mode, int epb, int nchunk) {
  cn_interior_body<NI>(a, mode, epb, nchunk, hw_blk());
}

int cn_interior_blocks(const CnArgs& a) {
  const int epb = std::min(32, kCnInteriorThreads / a.ni);
  const int nchunk = (a.nfe + epb - 1) / epb;
  return a.L1 * nchunk;
}

void launch_cn_stage(const CnArgs& a, const OneBodyElArgs* ob, int shape, int lc, int mode, hipStream_t s) {
  const int nd = a.nfun - a.nfcore;
  if (nd <= 0) return;
  if (mode == 3 || mode == 4) {  // y = zfac1 (T + V) x, or zfac1 V x (flag clear)
    if (active_recorder()) return record_unsupported("an application of h1");
    if (ob) {
      OneBodyElArgs e = *ob;
      e.cn = mode == 3 ? 3 : 2;
      e.cv = a.zfac1;
      e.stop = a.flag;
      launch_onebody_el(a.x, a.y, e, s, shape, lc);
    } else {
      hipLaunchKernelGGL(k_cn_rhs, dim3((a.nr + 255) / 256, a.L1, nd), dim3(256), 0, s, a, mode == 3 ? 3 : 1);
    }
    return;
  }
  // with the element right-hand side (stream launches) each term's
  // convergence test runs at the start of the next term's right-hand side
  // (the test at the top of h1itr_prop's next cycle) instead of as a launch
  // of its own; a batch records the test as a launch (its series word)
  const bool fused = ob != nullptr && !active_recorder();
  if (mode == 2) {  // a plain solve: y = zfac0 psi, then the first term
    if (active_recorder()) return record_unsupported("a rational-function solve");
    launch_zaxpby((size_t)nd * a.L1 * a.nr, a.zfac0, a.wfn + (size_t)a.nfcore * a.L1 * a.nr, make_double2(0.0, 0.0),
                  nullptr, a.y, s);
    mode = 0;
  } else if (ob) {  // element form of the right-hand side (ndvr = 20)
    OneBodyElArgs e = *ob;
    e.cn = mode + 1;
    if (mode == 1 && fused) {  // the previous term's convergence test first
      e.chk_part = a.part;
      e.chk_n = a.npart1 + a.L1 * ((a.nfe + std::min(32, kCnInteriorThreads / a.ni) - 1) /
                                   std::min(32, kCnInteriorThreads / a.ni));
      e.chk_nfun = a.nfun;
      e.chk_thresh = a.thresh;
      e.chk_flag = a.flag;
    }
    launch_onebody_el(mode == 0 ? a.wfn + (size_t)a.nfcore * a.L1 * a.nr : a.x, a.y, e, s, shape, lc);
  } else {
    if (active_recorder()) return record_unsupported("the band right-hand side of the half step (ndvr != 20)");
    hipLaunchKernelGGL(k_cn_rhs, dim3((a.nr + 255) / 256, a.L1, nd), dim3(256), 0, s, a, mode);
  }
  const int epb = std::min(32, kCnInteriorThreads / a.ni);
  const int nchunk = (a.nfe + epb - 1) / epb;
  const dim3 grid(a.L1 * nchunk);
  if (active_recorder()) {
    if (a.ni != 19) return record_unsupported("the half step with ndvr != 20");
    CnRec r;
    std::memset(&r, 0, sizeof r);
    r.a = a; r.mode = mode; r.epb = epb; r.nchunk = nchunk;
    const int var[4] = {mode, 0, 0, 0};
    record_launch(kRecCnBridge, var, dim3(nd * a.L1), 256, 0, r);
    record_launch(kRecCnInterior, var, grid, kCnInteriorThreads, 0, r);
    record_launch(kRecCnCheck, var, dim3(1), kCnInteriorThreads, 0, r);
    return;
  }
  hipLaunchKernelGGL((k_cn_bridge<256>), dim3(nd * a.L1), dim3(256), 0, s, a, mode);
  if (a.ni == 19)
    hipLaunchKernelGGL((k_cn_interior<19>), grid, dim3(kCnInteriorThreads), 0, s, a, mode, epb, nchunk);
  else
    hipLaunchKernelGGL((k_cn_interior<0>), grid, dim3(kCnInteriorThreads), 0, s, a, mode, epb, nchunk);
  if (!fused) hipLaunchKernelGGL(k_cn_check, dim3(1), dim3(kCnInteriorThreads), 0, s, a, mode, (int)grid.x);
}

void launch_cn_check(const CnArgs& a, hipStream_t s) {
  const int epb = std::min(32, kCnInteriorThreads / a.ni);
  const int nblk = a.L1 * ((a.nfe + epb - 1) / epb);
  hipLaunchKernelGGL(k_cn_check, dim3(1), dim3(kCnInteriorThreads), 0, s, a, 1, nblk);
}

__global__ void __launch_bounds__(256) k_cn_bridge_batch(const CnRec* __restrict__ recs, const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(CnRec)];
  int have = -1;
  // the first term opens the series' word (read by no launch of this term)
  if (blockIdx.x == 0 && threadIdx.x == 0 && recs[0].mode == 0 && recs[0].gate.word) *recs[0].gate.word = 0;
  batch_walk(map, [&](int m) { return recs[m].mode == 1 && (gate_closed(recs[m].gate) || recs[m].a.flag[0] != 0); },
             [&](int m, const Blk& bk) {
    const CnRec& r = stage_rec(recs, m, raw, have);
    cn_bridge_body<256>(r.a, r.mode, bk);
  });
}

__global__ void __launch_bounds__(kCnInteriorThreads) k_cn_interior_batch(const CnRec* __restrict__ recs,
                                                                          const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(CnRec)];
  int have = -1;
  batch_walk(map, [&](int m) { return recs[m].mode == 1 && (gate_closed(recs[m].gate) || recs[m].a.flag[0] != 0); },
             [&](int m, const Blk& bk) {
    const CnRec& r = stage_rec(recs, m, raw, have);
    cn_interior_body<19>(r.a, r.mode, r.epb, r.nchunk, bk);
  });
}

__global__ void __launch_bounds__(kCnInteriorThreads) k_cn_check_batch(const CnRec* __restrict__ recs,
                                                                       const BatchMap map) {
  __shared__ __align__(16) unsigned char raw[sizeof(CnRec)];
  int have = -1;
  batch_walk(map, [&](int m) { return recs[m].mode == 1 && recs[m].a.flag[0] != 0; }, [&](int m, const Blk& bk) {
    const CnRec& r = stage_rec(recs, m, raw, have);
    cn_check_body(r.a, r.mode, r.a.L1 * r.nchunk);
    // (no gate on this launch: its blocks write the word that the gate reads)
    if (threadIdx.x == 0 && r.gate.word && r.a.flag[0] == 0) *r.gate.word = r.gate.term;
  });
}

// One block per job (group g, l, element chunk) over the group's kCnGroup
// (member, orbital) items: cn_interior_body with the row of G_e loaded once
// for all items (the shared table holds every member's own numbers), each
// item's block norm reduced in the order of the member's own pass and
// written to the member's partial-sum slot of that (l, chunk) block.
// Items of a member whose series has converged (mode 1, flag set) or
// outside its extent are passed over.
__global__ void __launch_bounds__(kCnInteriorThreads)
    k_cn_interior_grouped(const CnRec* __restrict__ recs, const int2* __restrict__ items,
                          const int* __restrict__ jobs, const CnShared sh) {
  constexpr int S = kCnGroup, NI = 19, n = 20, BS = kCnInteriorThreads;
  if (gate_closed(sh.gate)) return;
  const int job = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int g = jobs[3 * job], l = jobs[3 * job + 1], ch = jobs[3 * job + 2];
  const int epb = sh.epb, e0 = ch * epb;
  const int t = threadIdx.x;
  const int el = t / NI, i = t - el * NI;
  const int e = e0 + el;
  const int r = e * n + i;
  __shared__ double2 sb[S][BS];
  __shared__ double2 sx[S][34];  // bridges e0-1 .. e0+epb-1 (epb <= 32)
  __shared__ double sRed[S][BS / 64];
  __shared__ int sstate[S], snfe[S], smode[S];  // state 0: pass over, 1: norm 0 only (l < |m|), 2: solve
  __shared__ const double2 *sy[S], *sWl[S], *sWr[S], *sxb[S];
  __shared__ double2 *sxo[S], *swf[S];
  __shared__ double* spart[S];
  if (t < S) {
    const int2 it = items[g * S + t];
    int st = 0;
    snfe[t] = 0;
    if (it.x >= 0) {
      const CnRec& rc = recs[it.x];
      const CnArgs& a = rc.a;
      if (!(rc.mode == 1 && a.flag[0]) && l < a.L1 && ch < rc.nchunk) {
        const int fd = it.y, m = a.mval[a.nfcore + fd];
        st = l >= (m < 0 ? -m : m) ? 2 : 1;
        const size_t off = ((size_t)fd * a.L1 + l) * a.nr;
        snfe[t] = a.nfe;
        smode[t] = rc.mode;
        sy[t] = a.y + off;
        sxo[t] = a.x + off;
        sxb[t] = a.x + off + (n - 1);  // bridge b at sxb[b * n]
        swf[t] = a.wfn + ((size_t)(a.nfcore + fd) * a.L1 + l) * a.nr;
        sWl[t] = a.Wl + (size_t)l * a.nr;
        sWr[t] = a.Wr + (size_t)l * a.nr;
        spart[t] = a.part + a.npart1 + xcd_unmap(l * rc.nchunk + ch, a.L1 * rc.nchunk);
      }
    }
    sstate[t] = st;
  }
  __syncthreads();
  const bool in_blk = el < epb && e < sh.gstride;
#pragma unroll
  for (int j = 0; j < S; ++j)
    if (sstate[j] == 2 && in_blk && e < snfe[j]) sb[j][t] = sy[j][r];
  for (int q = t; q < S * (epb + 1); q += BS) {
    const int j = q / (epb + 1), k = q - j * (epb + 1);
    const int bidx = e0 - 1 + k;
    sx[j][k] = (sstate[j] == 2 && bidx >= 0 && bidx < snfe[j] - 1) ? sxb[j][(size_t)bidx * n]
                                                                   : make_double2(0.0, 0.0);
  }
  double2 gk[NI];
  if (in_blk) {
    const double2* ge = sh.G + (((size_t)l * sh.gstride) + e) * NI * NI + i;
#pragma unroll
    for (int k = 0; k < NI; ++k) gk[k] = ge[k * NI];
  }
  __syncthreads();
  double nrm[S];
#pragma unroll
  for (int j = 0; j < S; ++j) nrm[j] = 0.0;
  if (in_blk) {
    const int eb = el * NI;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int nfe = snfe[j];
      if (sstate[j] != 2 || e >= nfe) continue;
      double2 x = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < NI; ++k) cfma(x, gk[k], sb[j][eb + k]);
      if (e >= 1) x = csub(x, cmul(sWl[j][r], sx[j][el]));
      if (e <= nfe - 2) x = csub(x, cmul(sWr[j][r], sx[j][el + 1]));
      sxo[j][r] = x;
      double2* w = swf[j] + r;
      *w = smode[j] == 0 ? x : cadd(*w, x);
      nrm[j] += x.x * x.x + x.y * x.y;
    }
  }
#pragma unroll
  for (int j = 0; j < S; ++j) {
    if (sstate[j] == 0) continue;
    double v = nrm[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((t & 63) == 0) sRed[j][t >> 6] = v;
  }
  __syncthreads();
  if (t < S && sstate[t] != 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) s += sRed[t][w];
    *spart[t] = s;
  }
}

void launch_cn_interior_grouped(const void* recs, const int2* items, const int* jobs, int njobs, const CnShared& sh,
                                hipStream_t s) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(k_cn_interior_grouped, dim3(njobs), dim3(kCnInteriorThreads), 0, s,
                     static_cast<const CnRec*>(recs), items, jobs, sh);
}

bool launch_batch_h1itr(const LaunchRec& p, const void* recs, const BatchMap& map, int grid, size_t lds,
                        hipStream_t s, bool dry) {
  const CnRec* r = static_cast<const CnRec*>(recs);
  if (p.kind == kRecCnBridge && p.block == 256) {
    if (!dry) hipLaunchKernelGGL(k_cn_bridge_batch, dim3(grid), dim3(256), lds, s, r, map);
  } else if (p.kind == kRecCnInterior && p.block == kCnInteriorThreads) {
    if (!dry) hipLaunchKernelGGL(k_cn_interior_batch, dim3(grid), dim3(kCnInteriorThreads), lds, s, r, map);
  } else if (p.kind == kRecCnCheck && p.block == kCnInteriorThreads) {
    if (!dry) hipLaunchKernelGGL(k_cn_check_batch, dim3(grid), dim3(kCnInteriorThreads), lds, s, r, map);
  } else {
    return false;
  }
  return true;
}

}  // namespace td1c
