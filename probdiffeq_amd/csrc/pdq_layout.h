// LDS layout of one instance in the cooperative kernel (offsets in doubles).
// Shared by the host launcher (to size dynamic LDS) and the device code.
#pragma once

#include "../../include/pdq.h"

namespace pdq {

struct Layout {
  int nc, mc, mcp, ko, N, d;
  int smoother, ts1, expo;
  // one solution slot
  int s_sc, s_m, s_L, s_cA, s_cm, s_cL, s_ctl, s_cto, s_b, s_J, s_size;
  // instance workspace (after the three slots)
  int w_P, w_PI, w_big, w_T1, w_Um, w_UL, w_Pm, w_PL, w_om, w_oL, w_G, w_mo, w_y0;
  int w_bA, w_bm, w_bL, w_btl, w_btl_pad, w_bto;
  int w_eA, w_em, w_eL, w_etl, w_eto, w_prm, w_tri, w_lml;
  int w_svd;  // the LML gain's singular value decomposition of R_Y (U, V, s)
  // exponential prior: the transition of the current step (e^A, Gramian
  // factor) and the scratch of its computation
  int w_xA, w_xL, w_xa, w_xB, w_xA2, w_xA4, w_xA6, w_xA8, w_xT, w_xV, w_xVmU, w_xP, w_xPB, w_xrhs, w_xtmp,
      w_xbuf, w_xpiv;
  int inst_size;   // doubles per instance
  int wg_shared;   // doubles shared by the workgroup (A, Q; the exponential prior's drift rows)
};

#if defined(__HIPCC__)
#define PDQ_HD __host__ __device__
#else
#define PDQ_HD
#endif

// n_tan: forward-mode tangent columns carried by every mean (0 = none).
PDQ_HD inline Layout make_layout(const pdq_config& cfg, int n_tan) {
  Layout L{};
  L.N = cfg.num_derivatives + 1;
  L.d = cfg.ode_dim;
  const bool dense = cfg.factorisation == PDQ_DENSE;
  L.nc = dense ? L.N * L.d : L.N;
  L.mcp = dense ? 1 : L.d;
  L.mc = L.mcp * (1 + n_tan);
  L.ko = dense ? L.d : 1;
  L.smoother = cfg.strategy != PDQ_FILTER;
  L.ts1 = cfg.constraint == PDQ_TS1;
  L.expo = dense && cfg.prior == PDQ_PRIOR_EXPONENTIAL;
  const int nc = L.nc, mc = L.mc, ko = L.ko;
  int o = 0;
  L.s_sc = o; o += 8;
  L.s_m = o; o += nc * mc;
  L.s_L = o; o += nc * nc;
  if (L.smoother) {
    L.s_cA = o; o += nc * nc;
    L.s_cm = o; o += nc * mc;
    L.s_cL = o; o += nc * nc;
    L.s_ctl = o; o += nc;
    L.s_cto = o; o += nc;
  } else {
    L.s_cA = L.s_cm = L.s_cL = L.s_ctl = L.s_cto = 0;
  }
  L.s_b = o; o += ko * mc;
  L.s_J = o; o += L.ts1 ? ko * nc : 1;
  L.s_size = o;
  o = 3 * L.s_size;
  const int big1 = 4 * nc * nc, big2 = (ko + nc) * (ko + nc);
  L.w_P = o; o += nc;
  L.w_PI = o; o += nc;
  L.w_big = o; o += big1 > big2 ? big1 : big2;
  L.w_T1 = o; o += nc * mc;
  L.w_Um = o; o += nc * mc;
  L.w_UL = o; o += nc * nc;
  L.w_Pm = o; o += nc * mc;
  L.w_PL = o; o += nc * nc;
  L.w_om = o; o += ko * mc;
  L.w_oL = o; o += ko * ko;
  L.w_G = o; o += nc * ko;
  L.w_mo = o; o += ko * mc;
  L.w_y0 = o; o += L.d * (1 + n_tan);
  L.w_bA = o; o += nc * nc;
  L.w_bm = o; o += nc * mc;
  L.w_bL = o; o += nc * nc;
  L.w_btl = o; o += nc;
  L.w_btl_pad = o;
  L.w_bto = o; o += nc;
  L.w_eA = o; o += nc * nc;
  L.w_em = o; o += nc * mc;
  L.w_eL = o; o += nc * nc;
  L.w_etl = o; o += nc;
  L.w_eto = o; o += nc;
  L.w_prm = o; o += cfg.n_params > 0 ? cfg.n_params : 1;
  L.w_tri = o; o += ko * mc;
  L.w_lml = o; o += 4 * (1 + n_tan) + 4 * ko;
  L.w_svd = o; o += 2 * ko * ko + ko;
  if (L.expo) {
    const int nn = nc * nc, rows = 10 * L.d > 2 * nc ? 10 * L.d : 2 * nc;
    L.w_xA = o; o += nn;
    L.w_xL = o; o += nn;
    L.w_xa = o; o += nn;
    L.w_xB = o; o += nc * L.d;
    L.w_xA2 = o; o += nn;
    L.w_xA4 = o; o += nn;
    L.w_xA6 = o; o += nn;
    L.w_xA8 = o; o += nn;
    L.w_xT = o; o += nn;
    L.w_xV = o; o += nn;
    L.w_xVmU = o; o += nn;
    L.w_xP = o; o += nn;
    L.w_xPB = o; o += nc * L.d;
    L.w_xrhs = o; o += nc * 10 * L.d;
    L.w_xtmp = o; o += nc * (nc > 5 * L.d ? nc : 5 * L.d);
    L.w_xbuf = o; o += rows * nc;
    L.w_xpiv = o; o += nc;
  } else {
    L.w_xA = L.w_xL = L.w_xa = L.w_xB = L.w_xA2 = L.w_xA4 = L.w_xA6 = L.w_xA8 = L.w_xT = L.w_xV = L.w_xVmU =
        L.w_xP = L.w_xPB = L.w_xrhs = L.w_xtmp = L.w_xbuf = L.w_xpiv = 0;
  }
  o += 16;  // scalars / padding
  L.inst_size = o;
  L.wg_shared = 2 * nc * nc + (L.expo ? L.d * nc : 0);
  return L;
}

}  // namespace pdq
