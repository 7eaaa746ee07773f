// Device-side data layout shared by the kernels and the host packer.
//
// A batch holds any number of spectrum requests.  Per-request arrays are
// concatenated into pooled device buffers; DevReq records the offsets.
// Work items:
//   layer kernel     one wavefront per (request, wavenumber, layer)
//   emission kernel  one wavefront per (request, wavenumber, path)
//   quadrature       one wavefront per (request, wavenumber)
#pragma once
#include <stdint.h>

#include "../../include/radg.h"

#define RADG_WAVE 64
#define RADG_FUSED_WAVES 16  // wavefronts per workgroup of the fused kernel
#define RADG_MAXPARAM 41  // ngas + 1 + ncont with ngas <= 30, ncont <= 10 (no para-H2)

// gas-continuum kinds (ngascon.f:144-262) and the per-wavenumber factors the
// host forms for them: CH4 and NH3 coefficients (1e20 cm2/molecule), the H-
// bound-free cross-section, and the CKD H2O terms (in range flag, self
// continuum at 296 and 260 K, foreign continuum with its CKD 2.1 scaling,
// the self continuum's scaling)
enum { RADG_GC_NONE = 0, RADG_GC_H2O = 1, RADG_GC_CH4 = 2, RADG_GC_NH3 = 3, RADG_GC_HBF = 4,
       RADG_GC_HFF = 5 };
enum { RADG_GCW_CH4 = 0, RADG_GCW_NH3, RADG_GCW_HBF, RADG_GCW_H2OIN, RADG_GCW_S296,
       RADG_GCW_S260, RADG_GCW_F296, RADG_GCW_SFAC, RADG_GCW };

struct DevReq {
  int L, ngas, npath, nparam, nsw, nwave;
  int ispace, tsurf_le0, inormal;
  int iray;                    // Rayleigh form 1..4, 0 = off (iray > 0 and itype != 0)
  int ncont;                   // dust types; nparam = ngas + 1 + ncont
  int inlte;                   // calc_nlte_t.f flag (0 = LTE)
  int flagh2p;                 // 1: para-H2 CIA (nparacon_all.f), fpara the last parameter
  int solar;                   // stellar spectrum in use: 1 divides imod 3/18, 4 scales imod 0
  long long dx0;               // per-(wavenumber, dust type) extinction
  long long ct0;               // per-(layer, dust type) particle columns
  double tsurf;
  double nlte_k1, nlte_a;      // calc_nlte_t.f constants
  double xfac;                 // imod 3 output scale (iform 1, 3: total flux)
  double xfac18;               // imod 18 output scale, its 2 pi included
  double solar_area;           // iform 4: 4 pi (dist AU)^2 (1 for dist <= 0)
  int iabs[5];                 // nciacon gas slots (0-based gas index, -1 = absent)
  // gas continua (ngascon.f): kind per gas (RADG_GC_*), the e- (80) and H-
  // (91) gas slots the H- free-free term reads, and the offset of the
  // per-wavenumber factors [nwave][RADG_GCW] (-1: no continuum gas)
  signed char gcon[RADG_MAXGAS];
  short ke80, ke91;
  long long gcw0;
  // gases with a k-table: row of each gas in the layer results' dk (-1: the
  // gas has no table and its dk/dq is zero), nks of them
  signed char kslot[RADG_MAXGAS];
  int nks;
  // dk row of each Jacobian parameter isw[k1] (kslot for a gas, nks for T,
  // -1 for a gas without a table and the continuum-only parameters)
  signed char kdrow[RADG_MAXPARAM];
  int isw[RADG_MAXPARAM];      // 0-based parameter ids
  long long lay0;              // LayerConst / per-layer arrays
  long long lg0;               // per-(layer, gas) arrays: amo
  long long path0;             // per-path arrays
  long long pl0;               // per-(path, layer) arrays
  long long wave0;             // per-wavenumber arrays
  long long wg0;               // per-(wavenumber, gas) slice records
  long long itemA0, itemB0, itemC0;
  long long inter0;            // intermediate buffer, elements
  long long out0;              // output[nwave][npath]
  long long jac0;              // jac[nwave][npath][L][nparam]
  long long wout0;             // wout[nwave]
  long long wjac0;             // wjac[nwave][L][nparam]
};

// Per-layer constants, computed on the host in FP64 (get_kg.f:113-174 for
// the table bracket, cirsradg_wave.f:418-426 for utotl, nciacon.f:65-176
// for the CIA mixing ratios and path factor).
struct LayerConst {
  int cp, ct;            // 1-based lower bracket indices in P and T
  double v, u, dudt;     // P weight, T weight, dU/dT
  // the T bracket on the upper pressure's grid: the same as ct, u, dudt for
  // one temperature grid, its own for per-pressure grids (NT < 0,
  // get_kg.f:145-173)
  int ct2, pad2;
  double u2, dudt2;
  double utotl;          // sum of amounts * 1e-20
  double temp;           // K
  double qh2, qhe, qn2, qch4, qco2, tau;
  double press;          // atm
  double fpara;          // para-H2 fraction (hfp), FLAGH2P = 1
  double fheh2, fch4h2, fnh3;  // rayleighls mixture, cirsradg_wave.f:520-523,671-677
};

// Table slice(s) used by one gas at one wavenumber.
struct WaveGas {
  long long base;        // element offset of the slice in the table pool; -1: k = 0
  long long base2;       // next slice up in wavenumber (non-coincident only)
  double w2;             // weight of the upper slice, get_kg.f:381-384
  // 16-bit forms: index of each slice's per-ordinate offsets in the offset
  // pool (log10 k = code + offset[g])
  long long ob, ob2;
  int coinc;             // 1: the wavenumber is on the table grid (get_kg.f:253-264)
  int pad;
};

// 16-bit codes of k <= 0 (-infinity): BF16 and IEEE half
#define RADG_BF16_NONPOS 0xFF80u
#define RADG_F16_NONPOS 0xFC00u

// nlays: the path's layers, and the Jacobian rows it writes; for an imod 8
// path (cell x atmosphere, cirsradg_wave.f:1037-1082) nlayin(1), the rows
// the reference sums and maps (:1251-1252, map2pro.f:52-57), and cell/atm
// the 0-based paths it multiplies
struct PathInfo {
  int nlays, imod, surf, cell, atm, pad[3];
};

// Intermediate block per (request, wavenumber), in elements of the compute
// type:  ktot[NG][L] | dk[NG][nks+1][L] | tcon[L] | dtcon[nparam][L]
// dk holds derivatives with respect to amount*1e-20 for the nks gases with a
// k-table (row kslot[gas]; the reference's dkdq before its 1e-20 factor,
// cirsradg_wave.f:749) and with respect to T in row nks.  dtcon holds them
// for every gas (slot = gas) and T (slot ngas); its further rows are the
// continuum's only parameters: each dust type (its extinction, the same in
// every layer) and the para-H2 fraction.
__host__ __device__ inline long long inter_block(int ng, int L, int nks, int nparam) {
  return (long long)ng * L * (nks + 2) + (long long)L * (1 + nparam);
}

// LDS of the layer pass.  A workgroup holds a header shared by its
// wavefronts, then one scratch area per wavefront for the layer item it is
// on.  Header: G [ng+6] (the g boundaries, G[ng] forced >= 1, then +inf so
// that a search may look five edges past the last), delg [ng] in FP64 and
// in the compute type, the FP64 weight of each of the 64*E sort slots, and
// Gw [ng+1], the unforced cumulative weights sum_{j<c} delg_j.
__host__ __device__ inline int lds_round16(int b) { return (b + 15) & ~15; }
struct HeaderLayout {
  int G, dl64, dlg, w64, Gw, total;
};
__host__ __device__ inline HeaderLayout header_layout(int ng, int E, int ts) {
  HeaderLayout H;
  int o = 0;
  H.G = o;    o += lds_round16(8 * (ng + 6));
  H.dl64 = o; o += lds_round16(8 * ng);
  H.dlg = o;  o += lds_round16(ts * ng);
  H.w64 = o;  o += 8 * RADG_WAVE * E;
  H.Gw = o;   o += lds_round16(8 * (ng + 1));
  H.total = o;
  return H;
}
__host__ __device__ inline int layer_header_bytes(int ng, int E, int ts) {
  return header_layout(ng, E, ts).total;
}
// Sort slots of an overlap round: slot s = l*E + q of lane l holds element
// (i, j) = (l / lpr, (l % lpr) * E + q), lpr = ceil(ng / E) lanes per row of
// the running distribution; slots past a row's ng elements, and lanes past
// ng * lpr, are padding.  s grows with (i, j) in lexicographic order.
__host__ __device__ inline int layer_lpr(int ng, int E) { return (ng + E - 1) / E; }
// Row-count rebin of an overlap round (E >= 4): the bin edges are shared
// by ngrp = 64 / ng groups of ng lanes, each group summing over about
// ng / ngrp rows of the running distribution.
__host__ __device__ inline int count_groups(int ng) { return RADG_WAVE / ng; }
// rows a lane sums over at most (E = 4: 13-16 g-ordinates, 8: 17-21, 16: 22-32)
__host__ __device__ inline int count_rpg(int E) { return E == 4 ? 4 : E == 8 ? 7 : 16; }
// rows of rrec: count_rpg(E) per group, group g's first rpg = ceil(ng /
// ngrp) of them rows g rpg + t of the distribution (while < ng), the rest zero
__host__ __device__ inline int count_rows(int ng, int E) { return count_groups(ng) * count_rpg(E); }
// Its scratch, inside `uni` (byte offsets from uni):
//   gam  [64 E]         bytes: the number of bin edges at or below the
//                       cumulative weight through each sort slot's element,
//                       indexed by the element's slot
//   ctab [ng+1][3]      sum_{j<c} delg_j, delg_j k2_j and delg_j dk2_j/dT (FP64)
//   part [ngrp-1][ng][5] the other groups' partial sums (FP64), over gam and
//                       ctab once the counts are summed
//   recI [ng]           sort slot of the element straddling each edge
//   rrec [nrow][ngas+3] per row i: delg_i, delg_i K_i, delg_i dK_i/dT, then
//                       delg_i times each gas payload (FP64), in group
//                       order with zero rows (nrow = count_rows); past every
//                       other use of uni (`keep`), so the weights and zero
//                       rows written once per item survive the rounds
struct CountLayout {
  int gam, ctab, part, recI, rrec, total;
};
__host__ __device__ inline CountLayout count_layout(int ng, int ngas, int E, int keep) {
  CountLayout C;
  C.gam = 0;
  C.part = 0;
  C.ctab = lds_round16(RADG_WAVE * E);
  const int tabs = C.ctab + lds_round16(24 * (ng + 1));
  const int part = lds_round16(40 * (count_groups(ng) - 1) * ng);
  int o = tabs > part ? tabs : part;
  C.recI = o; o += lds_round16(4 * ng);
  if (o < keep) o = keep;
  C.rrec = o; o += lds_round16(8 * count_rows(ng, E) * (ngas + 3));
  C.total = o;
  return C;
}
// the bytes of uni the other users need: the rankk merge buffers, the exact
// FP32 keys of a sort, the record rebin's records
__host__ __device__ inline int uni_keep(int ng, int ngas, int E, int ts) {
  int u = 2 * 8 * ng * (ngas + 2);
  const int merge = 3 * RADG_WAVE * ts + (ts == 4 ? 4 * RADG_WAVE : 0);
  if (u < merge) u = merge;
  if (ts == 4 && u < 4 * RADG_WAVE * E) u = 4 * RADG_WAVE * E;
  return lds_round16(u);
}
// Byte offsets within one wavefront's scratch (ts: bytes of the compute
// type).  `uni` is shared, in time, by the rankk merge buffers, the exact
// FP32 keys of an overlap sort and its rebin's records or row-count tables.
struct LayerLayout {
  int sSt, sGd, kgn, dkgn, bufA, bufB, sE, uni, total;
};
__host__ __device__ inline LayerLayout layer_layout(int ng, int ngas, int E, int ts) {
  LayerLayout L;
  int o = 0;
  L.sSt = o;  o += lds_round16(8 * ng);
  L.sGd = o;  o += lds_round16(8 * ng);
  L.kgn = o;  o += lds_round16(ts * ngas * ng);
  L.dkgn = o; o += lds_round16(ts * ngas * ng);
  L.bufA = o; o += lds_round16(ts * (ngas + 2) * ng);
  L.bufB = o; o += lds_round16(ts * (ngas + 2) * ng);
  L.sE = o;   o += lds_round16(4 * ng);
  int u = uni_keep(ng, ngas, E, ts);
  if (E >= 4) u = count_layout(ng, ngas, E, u).total;
  L.uni = o;  o += lds_round16(u);
  L.total = o;
  return L;
}

struct KernelArgs {
  const DevReq *reqs;
  int nreq;
  const long long *startA, *startB, *startC;  // [nreq+1] item prefix sums
  long long nA, nB, nC;
  // launches run one class of requests at a time: startA/nA (layer and
  // continuum items) and startB/nB (emission items, or wavenumber items of
  // the fused kernel) then cover the class's requests, listed in
  // reqmapA[nreqA] and reqmapB[nreqB]
  const int *reqmapA, *reqmapB;
  int nreqA, nreqB;
  int wsbytes;                 // fused kernel: LDS scratch bytes per wavefront
  int lwaves, lwstride;        // layer kernel: wavefronts (items) per workgroup, scratch bytes each
  int ng;
  int rebin;                   // overlap rebin: 1 row counts where they apply, 0 records only
  int lbl;                     // 1: line-by-line tables (ng = 1), the LBL layer kernel
  // tables
  const void *ktab;            // double* (FP64), float* (mixed) or 16-bit codes
  const float *koff;           // 16-bit forms: offset per (slice, ordinate)
  int np, nt;                  // nt: temperatures per pressure (|NT|)
  int t2;                      // 1: a temperature grid per pressure (NT < 0)
  const double *G;             // g boundaries [ng+1], G[ng] forced >= 1
  const double *delg64;
  const float *delg32;
  const char *hdr_img;         // the layer kernel's LDS header, built once per table set and form
  // CIA
  const double *kcia;          // [1501][25][9]
  const double *cia_temps;     // [25]
  double cia_dnu;
  // requests
  const LayerConst *lay;
  const double *amo;           // [layer][gas] amount*1e-20
  const double *press;         // [layer]
  const PathInfo *paths;
  const int *layinc;           // [path][L] 0-based
  const double *scale, *emtemp;
  const double *vwave, *esurf;
  const double *n2nir;  // [wavenumber][2]: N2-N2, N2-H2 near-infrared CIA
  const double *solar;  // [wavenumber]: the stellar spectrum (iread = 999)
  const double *ppg;    // [layer][gas] partial pressures (atm), gas continua
  const double *gcw;    // [wavenumber][RADG_GCW] gas-continuum factors
  double h2o_c2, h2o_ts, h2o_ps;  // /COMCON/ for the CKD H2O continuum
  const WaveGas *wg;
  const double *pweight;       // [path]
  const double *dext;          // [wavenumber][ncont] dust extinction (cm2) at vwave
  const double *cont;          // [layer][ncont] dust particles/cm2
  // /FPTABLE/ (FLAGH2P = 1): kh2h2/kh2he [1501][25][npara], kcia5 [1501][25][5]
  const double *pkh2h2, *pkh2he, *pkcia5, *ptemps, *pfp;
  double pdnu;
  int npara;
  // /E3TABLE/ (imod 18): transmission, E3, gradient [3][RADG_E3N]
  const double *e3tab;
  // outputs
  void *inter;                 // intermediates, compute type
  void *out, *gts, *jac;       // compute type; jac gas columns per 1e20 molecules
  double *wout, *wgts, *wjac;
  int *err;                    // device error flag
};

// One product of a batched FP64 GEMM (radg_oe.hip):
//   C(m, n) (+)= alpha * sum_k A(m, k) * B(k, n) * (scaleK ? scaleK[k] : 1)
// with A(m, k) = A[(rowmapA ? rowmapA[m] : m) * sAm + k * sAk] (float or
// double, fixed per launch), B(k, n) = B[k * sBk + n * sBn] and
// C(m, n) = C[m * sCm + n * sCn].  lower_only skips output tiles that lie
// wholly above the diagonal.
struct GemmProblem {
  const void *A;
  const int *rowmapA;
  const double *B;
  const double *scaleK;
  double *C;
  long long sAm, sAk, sBk, sBn, sCm, sCn;
  int M, N, K, accumulate, lower_only, pad;
  // A holds a stored per-path Jacobian: along k (period kper = nparam), the
  // first kgas entries are gas columns per 1e20 molecules and are scaled by
  // 1e-20 on load.  kper = 0: no scaling.
  int kper, kgas;
  double alpha;
};
