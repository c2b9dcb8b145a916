// Fused passes over spatially binned events for gfx950: one workgroup per
// chunk of a bin, the chunk's image window held in LDS.
//
// Reference semantics: Reconstruction::bb_update
// (src/2d/strip/reconstruction.h:220-298) per event, Jacobi over events
// (reconstruction.h:121-149).  One wave handles one event at a time; its
// lanes take the event's box rows, 64 at a time, the 64 rows around the
// ellipse centre first (the longest rows share a group, so the short ones
// do not pad it).  Each lane
//   1. finds its row's exact ellipse column interval: the roots of the
//      row's quadratic give the two boundary columns to within a small
//      fraction of a pixel, and the reference's own float test
//      (in_ellipse, gaussian_kernel.h:131-137) decides each boundary column
//      and its outer neighbour; rows the estimate cannot settle take a
//      scanning path;
//   2. pass 1: walks the interval summing K * rho read from the LDS window;
//      the DPP wave sum is the denominator lambda_e (the forward projection);
//   3. pass 2: walks it again adding K * w_e (* rho) into the LDS
//      accumulation window as 64-bit fixed point.
// Modes:
//   MLEM      pass 1 + pass 2 with w_e = 1/lambda_e; each value K rho /
//             lambda_e <= 1 is added in 2^-31 units, or in 2^-40 units
//             (split conversion) for the deterministic accumulation
//   MLEM_K    MLEM with pass-2 values K / lambda_e, as GRAD adds them (no
//             rho read or multiply per pixel); rho multiplies the sums
//             once per pixel afterwards
//   FORWARD   pass 1 only, plus the sigma derivative of lambda_e and the
//             log-likelihood terms, reduced per workgroup into doubles
//   GRAD      FORWARD and a pass 2 with w_e = 1/lambda_e (no rho): the
//             log-likelihood gradient in one launch, values in 2^-k units
//             with a per-row-group choice of conversion
//   BACKWARD  pass 2 only with upstream weights g_e, values |K g_e| in
//             units 2^-k chosen on the host, added or subtracted by the
//             sign of g_e: each value below 2^31 when the largest |g_e| K
//             is within a factor 256 of the mean, else (split) each below
//             2^62 / n_events
// A value x becomes the integer round(x) with one fused multiply-add of 0.5
// and one float->u32 conversion; the event weight is folded into the
// exponent of K (below), so the pixel step of pass 2 is four VALU
// operations plus the LDS read and the LDS atomic.  The inner walks are
// unrolled by four so that each lane keeps several LDS reads and
// exponentials in flight.  The LDS windows are column-major with a column
// stride that is a multiple of 64, so the lanes (rows) of a wave always hit
// 64 different LDS banks.  At the end the window is
// added to the global accumulator with one atomic per touched pixel.
//
// Kernel value (GaussianKernel::operator(), gaussian_kernel.h:51-79): the
// exponent splits into a row part and a pixel part,
//   K = c/sqrt(a^T S a) exp(-isdl br^2/2) * exp(-isz pq^2)
//     = exp2(lsc_row - u^2),   u = pq sqrt(isz log2 e),
// with pq = dz - dy tan computed exactly as the reference at the row's first
// ellipse pixel and advanced by one pixel width per column (u = u0 + k du).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "strip_fused.h"
#include "strip_geometry.h"
#include "wave_ops.h"

namespace petstrip {

namespace {

__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Wave64 maximum of a non-negative int (same DPP pattern as wave_sum).
__device__ __forceinline__ int wave_max_i(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x124, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x128, 0xf, 0xf, false));
  int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
  int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
  return max(max(r0, r1), max(r2, r3));
}

__device__ __forceinline__ float wave_max_f(float v) {
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Exact reference membership of column iz in the row with (Ady2, Bdy).
__device__ __forceinline__ bool exact_in(const StripConst& s, const EventGeom& g,
                                         float Ady2, float Bdy, int iz) {
  float dz = fmaf((float)iz, s.pw, s.tl_z_half_w) - g.z;
  float q = fmaf(g.C * dz, dz, fmaf(Bdy, dz, Ady2));
  return q <= 9.0f;
}

constexpr float kLog2e = 1.44269504088896341f;

// Per-event values every row of the event uses.
struct EventRows {
  float zc;    // ellipse centre in column coordinates, (z - tl_z_half_w) / pw
  float kcol;  // 1 / (2 C pw): root offset in metres -> columns
  float c4;    // 4 C
  float c005;  // 4 C * 0.05
};

__device__ __forceinline__ EventRows event_rows(const StripConst& sg, const EventGeom& g,
                                               float inv_pw) {
  EventRows ev;
  ev.c4 = 4.0f * g.C;
  ev.c005 = ev.c4 * 0.05f;
  ev.kcol = 0.5f * inv_pw * __builtin_amdgcn_rcpf(g.C);
  ev.zc = (g.z - sg.tl_z_half_w) * inv_pw;
  return ev;
}

constexpr int kUnroll = 4;
constexpr int kSpareCols = kUnroll;  // zero rho window columns before the window

// One box row of an event as a lane walks it.
struct RowSeg {
  int len;        // ellipse pixels in the row (0: nothing to do)
  int w0;         // LDS window index of the first ellipse pixel
  float u0;       // u of the first ellipse pixel
  float lsc;      // log2 of the row factor
  float b1;       // 2^(-2 u du) at the walk's first pixel: the chain's ratio
  float dz_row;   // row part of d lnK / d sigma_z (FORWARD)
  float ddl_row;  // d lnK / d sigma_dl (FORWARD)
  int mg;         // groups of four the wave walks: its lanes' largest ceil(len / 4)
};

// The wave's walk length for its lanes' rows s (wave-uniform).
__device__ __forceinline__ void set_walk_groups(RowSeg& s) {
  s.mg = uniform(wave_max_i((s.len + 3) >> 2));
}

// The analytic kernel (AnalyticKernel::operator(), analytic_kernel.h:28-62)
// along a box row: with pq = pq0 + k pw at the walk's k-th pixel,
//   norm = a S a + 2 o S b = n0 + n1 pq,  b S a = a1 pq + b1,
//   b S b = 2 isz pq^2 + c1,  K = c / sqrt(norm) exp(-(bSb - (bSa)^2/norm)/2),
// with the row's vectors a, o and br as the per-event kernel forms them
// (strip_kernels.hip, analytic_value).  The exponent is rational in the
// column, so there is no geometric chain: every pixel takes a reciprocal
// square root, a reciprocal and an exponential.
struct RowA {
  float pq0;  // pq at the row's first ellipse pixel
  float n0, n1, a1, b1, c1;
};

__device__ __forceinline__ RowA analytic_row(const StripConst& sk, const EventGeom& g,
                                             float dy, float dz) {
  RowA r;
  const float sec_sq = g.sec * g.sec;
  const float yp = dy + g.y;
  const float ap = -(yp - sk.R) * sec_sq, aq = -(yp + sk.R) * sec_sq;
  const float ar = -2.0f * yp * g.sec * g.t;
  const float op = ap * g.t, oq = aq * g.t;
  const float orr = -yp * g.sec * (1.0f + 2.0f * g.t * g.t);
  const float br = -2.0f * dy * g.sec;
  const float aia = ap * sk.isz * ap + aq * sk.isz * aq + ar * sk.isdl * ar;
  r.a1 = sk.isz * ap + sk.isz * aq;
  r.b1 = br * sk.isdl * ar;
  r.c1 = br * sk.isdl * br;
  r.n0 = aia + 2.0f * (orr * sk.isdl * br);
  r.n1 = 2.0f * (sk.isz * op + sk.isz * oq);
  r.pq0 = dz - dy * g.t;
  return r;
}

// K at pq; bq = 2 isz.
__device__ __forceinline__ float analytic_at(float c, float bq, const RowA& r, float pq) {
  const float norm = fmaf(r.n1, pq, r.n0);
  const float bia = fmaf(r.a1, pq, r.b1);
  const float bib = fmaf(bq * pq, pq, r.c1);
  const float e = bib - bia * bia * __builtin_amdgcn_rcpf(norm);
  return c * __builtin_amdgcn_rsqf(norm) * __builtin_amdgcn_exp2f(e * (-0.5f * 1.44269504088896341f));
}

// The first box row of an event's row group 0 (which takes the 64 rows
// around the ellipse centre; later groups the rows above and below it, in
// order) for a box of nrows rows from row ry0.
__device__ __forceinline__ int group0_first_row(const StripConst& sg, const EventGeom& g,
                                                int ry0, int nrows, float inv_ph) {
  const int mid = (int)((sg.tl_y - g.y) * inv_ph) - ry0;
  return nrows > 64 ? min(max(mid - 32, 0), nrows - 64) : 0;
}

// The row's ellipse columns [a, b] under the reference's float test
// (in_ellipse, gaussian_kernel.h:131-137), within the event's box; false for
// an empty row.  dy: the row centre's offset from the ellipse centre.
__device__ __forceinline__ bool row_interval(const StripConst& sg, const EventGeom& g,
                                             const EventRows& ev, float dy, bool scan,
                                             int& a_out, int& b_out) {
  const float Ady2 = (dy * g.A) * dy;
  const float Bdy = dy * g.B;
  // C dz^2 + Bdy dz + Ady2 - 9: discriminant at the threshold; at 9.05 it
  // bounds what the float test can accept (its rounding is ~1e-6)
  const float d9 = fmaf(Bdy, Bdy, -ev.c4 * (Ady2 - 9.0f));
  if (!(d9 + ev.c005 >= 0.0f)) return false;
  const float sq = __builtin_amdgcn_sqrtf(fmaxf(d9, 0.0f));
  const float clo = fmaf(-(Bdy + sq), ev.kcol, ev.zc);  // first column, estimated
  const float chi = fmaf(sq - Bdy, ev.kcol, ev.zc);     // last column, estimated
  const int r = (int)floorf(clo + 0.5f);
  const int q = (int)floorf(chi + 0.5f);
  // Rows are hole-free intervals under the float test (checked on the host
  // from the constants), so a boundary is settled once a column and its
  // outer neighbour disagree.
  const bool in_r = exact_in(sg, g, Ady2, Bdy, r);
  const bool in_q = exact_in(sg, g, Ady2, Bdy, q);
  const bool in_r2 = exact_in(sg, g, Ady2, Bdy, in_r ? r - 1 : r + 1);
  const bool in_q2 = exact_in(sg, g, Ady2, Bdy, in_q ? q + 1 : q - 1);
  int a = in_r ? r : r + 1;
  int b = in_q ? q : q - 1;
  const bool left = in_r != in_r2, right = in_q != in_q2;
  // no column of [r-1, q+1] passes: the row is empty when the estimated
  // interval is shorter than two columns (every column in it was tested)
  const bool none = !(in_r || in_r2 || in_q || in_q2) && (chi - clo < 2.0f);
  if (scan || !((left && right) || none)) {
    // the estimate did not settle the row (or the caller asks for the
    // scan): scan a superset (roots at 9.05 widened by a column) from both
    // ends
    const float sq2 = __builtin_amdgcn_sqrtf(d9 + ev.c005);
    const float flo = floorf(fmaf(-(Bdy + sq2), ev.kcol, ev.zc)) - 1.0f;
    const float fhi = ceilf(fmaf(sq2 - Bdy, ev.kcol, ev.zc)) + 1.0f;
    a = (flo > (float)g.x0) ? (int)flo : g.x0;
    b = (fhi < (float)(g.x1 - 1)) ? (int)fhi : g.x1 - 1;
#pragma nounroll
    while (a <= b && !exact_in(sg, g, Ady2, Bdy, a)) ++a;
#pragma nounroll
    while (b >= a && !exact_in(sg, g, Ady2, Bdy, b)) --b;
  } else if (none) {
    return false;
  }
  a_out = max(a, g.x0);
  b_out = min(b, g.x1 - 1);
  return a_out <= b_out;
}

// Packed row interval of the precomputed table: (columns << 16) | (first
// column - the box's first column); 0 for an empty row.
__device__ __forceinline__ uint32_t pack_interval(const EventGeom& g, bool any, int a, int b) {
  return any ? ((uint32_t)(b - a + 1) << 16) | (uint32_t)(a - g.x0) : 0u;
}

// The chain's per-pixel ratio b1 = 2^(-2 u du) at the first pixel of the
// walk, which starts pad = 4 ceil(len / 4) - len pixels before the row.
__device__ __forceinline__ float row_ratio(float u0, int len, float du) {
  const int pad = ((len + kUnroll - 1) & ~(kUnroll - 1)) - len;
  const float u = fmaf(-(float)pad, du, u0);
  return __builtin_amdgcn_exp2f(-2.0f * u * du);
}

// The row's walk: its ellipse columns (clipped to the work item's [cx0,
// cx1)), u0 and the row factor at the first one, the chain ratio b1 and,
// SIGMA, the row's sigma derivatives; ANALYTIC: also the row's
// analytic-kernel constants into *ra.  use_tab (wave-uniform): t is the
// row's entry of the row table, whose columns replace the boundary search
// and, with tab_kernel (entries built for sk, s_pix and du; unclipped rows),
// whose u0, lsc and b1 are the row's.
template <bool SIGMA, bool ANALYTIC = false>
__device__ __forceinline__ RowSeg row_setup(const StripConst& sg, const StripConst& sk,
                                            const EventGeom& g, const EventRows& ev,
                                            int iy, int row_w, int cstride, float s_pix,
                                            float du, bool scan, int cx0, int cx1,
                                            RowA* ra = nullptr, bool use_tab = false,
                                            const RowEntry& t = RowEntry{0u, 0.0f, 0.0f, 0.0f},
                                            bool tab_kernel = false) {
  RowSeg s{0, 0, 0.0f, -1e30f, 1.0f, 0.0f, 0.0f};
  if constexpr (ANALYTIC) *ra = RowA{0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const float dy = fmaf(-(float)iy, sg.ph, sg.tl_y_half_h) - g.y;
  int a, b;
  if (use_tab) {
    if (t.iv == 0u) return s;
    a = g.x0 + (int)(t.iv & 0xffffu);
    b = a + (int)(t.iv >> 16) - 1;
  } else if (!row_interval(sg, g, ev, dy, scan, a, b)) {
    return s;
  }
  // the work item's columns [cx0, cx1) in the column-banded form
  a = max(a, cx0);
  b = min(b, cx1 - 1);
  if (a > b) return s;
  s.len = b - a + 1;
  s.w0 = row_w + a * cstride;
  if (!SIGMA && !ANALYTIC && use_tab && tab_kernel) {
    s.u0 = t.u0;
    s.lsc = t.lsc;
    s.b1 = t.b1;
    return s;
  }
  // pq at the first pixel, with the reference's rounding (gaussian_kernel.h:64)
  const float dz = fmaf((float)a, sg.pw, sg.tl_z_half_w) - g.z;
  s.u0 = fmaf(-dy, g.t, dz) * s_pix;
  if constexpr (ANALYTIC) *ra = analytic_row(sk, g, dy, dz);
  // kernel row factor (gaussian_kernel.h:58-78)
  const float sec_sq = g.sec * g.sec;
  const float yp = dy + g.y;
  const float ap = -(yp - sk.R) * sec_sq;
  const float aq = -(sk.R + yp) * sec_sq;
  const float ar = ((yp * -2.0f) * g.sec) * g.t;
  const float br = (dy * -2.0f) * g.sec;
  const float norm = fmaf(ar * sk.isdl, ar, fmaf(ap * sk.isz, ap, (aq * sk.isz) * aq));
  const float brb = (br * sk.isdl) * br;
  s.lsc = __builtin_amdgcn_logf(sk.c) - 0.5f * __builtin_amdgcn_logf(norm) -
          (0.5f * kLog2e) * brb;
  if constexpr (!ANALYTIC) s.b1 = row_ratio(s.u0, s.len, du);
  if constexpr (SIGMA) {
    // ln K = ln c - ln(norm)/2 - bSb/2 with c ~ 1/(sz^2 sdl), norm and bSb
    // linear in 1/sz^2 and 1/sdl^2:
    //   d lnK/d sz  = (-2 + isz (ap^2 + aq^2)/norm + 2 isz pq^2) / sz
    //   d lnK/d sdl = (-1 + isdl ar^2/norm + isdl br^2) / sdl
    const float apq = fmaf(ap, ap, aq * aq);
    const float rn = __builtin_amdgcn_rcpf(norm);
    s.dz_row = fmaf(sk.isz * apq, rn, -2.0f) * __builtin_amdgcn_rcpf(sk.sigma_z);
    s.ddl_row = (fmaf(sk.isdl * (ar * ar), rn, -1.0f) + brb) *
                __builtin_amdgcn_rcpf(sk.sigma_dl);
  }
  return s;
}

// k + sgn * log2(x) for x > 0 and integer k, formed as (k + sgn e) + sgn
// log2(m) with x = m 2^e: scaling x by a power of two while k moves the
// other way gives the same float, so pass-2 values scale exactly.
__device__ __forceinline__ float log2_shift(float k, float x, float sgn) {
  const int e = __builtin_amdgcn_frexp_expf(x);
  const float m = __builtin_amdgcn_frexp_mantf(x);
  return fmaf(sgn, (float)e, k) + sgn * __builtin_amdgcn_logf(m);
}

// Kernel values along a row.  With u = u0 + k du at the walk's k-th pixel,
//   K(k) = 2^(lsc - u^2) = 2^(lsc - u0^2 + e1 k) G(k),
//   e1 = -2 u0 du,  G(k) = 2^(-(k du)^2),
// the first factor is a geometric sequence in k for each lane and G(k) is
// the same for every lane, read from a table through the scalar cache: the
// walks take two multiplications per pixel instead of an exponential.  Each
// of the four unrolled pixels keeps its own chain, advanced by b4 = 2^(4 e1)
// per group; over a row's dozen groups the chain's relative drift stays
// below 1e-6.  The event multiplier of pass 2 (and the 2^-32 of the split
// conversion) enters as e_add.
struct Chain {
  float p[kUnroll];
  float b4;
};

// u0: u at the walk's first pixel; b1: the per-pixel ratio (row_ratio).
__device__ __forceinline__ Chain chain_start(float u0, float lsc, float b1, float e_add) {
  // one exponential, the first pixel's value; the other starts and b4 =
  // b1^4 by multiplication
  Chain c;
  const float e0 = fmaf(-u0, u0, lsc) + e_add;
  c.p[0] = __builtin_amdgcn_exp2f(e0);
#pragma unroll
  for (int u = 1; u < kUnroll; ++u) c.p[u] = c.p[u - 1] * b1;
  const float b2 = b1 * b1;
  c.b4 = b2 * b2;
  return c;
}

using ConstF = const __attribute__((address_space(4))) float;

// The lane's four LDS reads of a group (two ds_read2st64).
template <int HP>
__device__ __forceinline__ void read4(const float* rp, float (&rv)[kUnroll]) {
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) rv[u] = rp[u * HP];
}

// Pass 1 over one row group.  A lane walks its row in ng = ceil(len / 4)
// groups of four pixels that end at the row's last pixel: the first group
// starts pad = 4 ng - len (0 to 3) pixels before the row, and those pixels
// are dropped by zeroing their rho values (the window has spare zero columns
// in front, so the reads stay inside it).  Every kernel value then comes from
// the chain.  A lane whose groups are done re-reads its last one (so no read
// leaves the window) with its chain set to zero, which drops those pixels
// from every sum without a mask.
// s0 += sum K rho (two_sum error into comp when COMP); SIGMA: s1 += sum K
// rho u^2, from the moments M1 = sum j K rho and M2 = sum j^2 K rho of the
// walk index j, combined per row as u0^2 S + 2 u0 du M1 + du^2 M2 with u0 at
// the walk's first pixel.  The chain value times rho is weighted by G(j),
// j G(j) and j^2 G(j), all three from tables through the scalar cache: three
// FMAs per pixel give S, M1 and M2.
template <int HP, bool COMP, bool SIGMA, int UNR>
__device__ __forceinline__ void forward_rows(const RowSeg& s, float du, ConstF* gtab,
                                             ConstF* mtab, const float* rho0, float& s0,
                                             float& comp, float& s1) {
  const int ng = (s.len + kUnroll - 1) >> 2;
  const int pad = ng * kUnroll - s.len;
  const float u0 = fmaf(-(float)pad, du, s.u0);
  // an empty row reads the front spare columns
  const float* rp = s.len > 0 ? rho0 + (s.w0 - pad * HP) : rho0 - kSpareCols * HP;
  const int lastg = max(ng - 1, 0);
  const int maxng = s.mg;
  float sum0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
  Chain c = chain_start(u0, s.lsc, s.b1, 0.0f);
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) c.p[u] = ng > 0 ? c.p[u] : 0.0f;
  auto group = [&](int gk, bool first) {
    float rv[kUnroll];
    read4<HP>(rp + min(gk, lastg) * (kUnroll * HP), rv);
    if (first) {
#pragma unroll
      for (int u = 0; u < kUnroll - 1; ++u) rv[u] = u < pad ? 0.0f : rv[u];
    }
    const float b4 = gk + 1 < ng ? c.b4 : 0.0f;
    // COMP sums the group before its two-sum; otherwise every pixel goes
    // straight into the row sum
    float t0 = 0.0f;
    float& acc = COMP ? t0 : sum0;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = gk * kUnroll + u;
      if constexpr (SIGMA) {
        const float pr = c.p[u] * rv[u];
        acc = fmaf(pr, gtab[k], acc);
        m1 = fmaf(pr, mtab[2 * k], m1);
        m2 = fmaf(pr, mtab[2 * k + 1], m2);
      } else {
        acc = fmaf(c.p[u] * gtab[k], rv[u], acc);
      }
      c.p[u] *= b4;
    }
    if constexpr (COMP) {
      float sum, err;
      two_sum(sum0, t0, sum, err);
      sum0 = sum;
      comp += err;
    }
  };
  if (maxng > 0) group(0, true);
  // unrolled UNR times: one scalar load and one wait serve UNR groups
#pragma unroll UNR
  for (int gk = 1; gk < maxng; ++gk) group(gk, false);
  s0 += sum0;
  if constexpr (SIGMA) {
    const float a1 = 2.0f * u0 * du, a2 = du * du;
    s1 += fmaf(a2, m2, fmaf(a1, m1, u0 * u0 * sum0));
  }
}

// Pass 1 over one row group by Horner's rule (the denominator alone).  With
// the chain's values (chain_start) the walk's k-th pixel has
// K = P0 r^k G(k), r = b1 the chain's per-pixel ratio and P0 = 2^(lsc - us^2)
// (us: u at the walk's first pixel), so the row sum is
// P0 sum_j r^j p_j(R) over the group's four pixels j, with
// p_j(R) = sum_g R^g a_gj, R = r^4 and a_gj = G(4g + j) rho(4g + j): four
// polynomials in R, each evaluated by Horner's rule from the row's last group
// down to its first.  A pixel then takes the table weight and one FMA (the
// chain took a third operation to advance it), and the per-group chain
// zeroing and read clamping go: every lane walks groups G = mg - 1 down to 0
// in step (G is wave-uniform, so G(4G + j) comes through the scalar cache), a
// lane whose row has fewer groups reads zero columns of the window (zp) until
// its last group comes up, so its polynomials stay zero, and group 0's pixels
// before the row are zeroed.  The values R^g a_gj r^j are those of K rho / P0,
// as in the chain.  zp: the window's zero columns at the lane's row.
template <int HP, int UNR>
__device__ __forceinline__ void forward_rows_h(const RowSeg& s, float du, ConstF* gtab,
                                               const float* rho0, const float* zp, float& s0) {
  const int ng = (s.len + kUnroll - 1) >> 2;
  const int pad = ng * kUnroll - s.len;
  const float r = s.b1;  // the ratio at the walk's first pixel (row_ratio)
  const float r2 = r * r;
  const float R = r2 * r2;
  const float* base = rho0 + (s.w0 - pad * HP);
  float h[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) h[u] = 0.0f;
  auto step = [&](int G, const float(&rv)[kUnroll]) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) h[u] = fmaf(h[u], R, gtab[G * kUnroll + u] * rv[u]);
  };
  const int mg = s.mg;
#pragma unroll UNR
  for (int G = mg - 1; G > 0; --G) {
    float rv[kUnroll];
    read4<HP>(G < ng ? base + G * (kUnroll * HP) : zp, rv);
    step(G, rv);
  }
  if (mg > 0) {
    float rv[kUnroll];
    read4<HP>(ng > 0 ? base : zp, rv);
#pragma unroll
    for (int u = 0; u < kUnroll - 1; ++u) rv[u] = u < pad ? 0.0f : rv[u];
    step(0, rv);
  }
  if (s.len <= 0) return;
  const float us = fmaf(-(float)pad, du, s.u0);
  const float p0 = __builtin_amdgcn_exp2f(fmaf(-us, us, s.lsc));
  s0 += p0 * fmaf(fmaf(fmaf(h[3], r, h[2]), r, h[1]), r, h[0]);
}

// x -> the integer round(x) as (lo, hi); SPLIT: kv = x 2^-32 < 2^30 and the
// halves are converted apart; RPI: x < 2^31 is rounded by the conversion
// itself (v_cvt_rpi_i32_f32, floor(x + 0.5)); else x + 0.5 < 2^32 arrives
// with its 0.5 added and is truncated.
template <bool SPLIT, bool RPI = false>
__device__ __forceinline__ void to_fixed(float x_or_kv, unsigned& lo, unsigned& hi) {
  if constexpr (SPLIT) {
    hi = (unsigned)x_or_kv;
    lo = (unsigned)fmaf(__builtin_amdgcn_fractf(x_or_kv), 4294967296.0f, 0.5f);
  } else if constexpr (RPI) {
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x_or_kv));
    lo = (unsigned)r;
    hi = 0u;
  } else {
    lo = (unsigned)x_or_kv;  // x in [0.5, 2^32)
    hi = 0u;
  }
  // keep the values (and their LDS reads) ahead of the masked atomics
  // instead of letting each sink into its own branch
  asm volatile("" : "+v"(lo));
}

// One pass-2 value from the kernel value kv (which carries the event
// multiplier): K w rho or K w, + 0.5 for rounding unless SPLIT (whose
// conversion rounds the low half itself).
template <bool RHO, bool SPLIT, bool RPI = false>
__device__ __forceinline__ float fixed_arg(float kv, float rho) {
  if constexpr (SPLIT || RPI) return RHO ? kv * rho : kv;
  else if constexpr (RHO) return fmaf(kv, rho, 0.5f);
  else return kv + 0.5f;
}

template <bool SUB>
__device__ __forceinline__ void lds_accumulate(unsigned long long* p, unsigned long long v) {
  if constexpr (SUB)
    __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The 32-bit windows of the K / lambda sums (ds_add_u32, 1.5x the rate of
// ds_add_u64 on gfx950, one LDS bank per lane).
__device__ __forceinline__ void lds_accumulate32(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// (lo, hi) into a 64-bit register pair that lives across the loop: when hi
// is the constant 0 the pair's upper register keeps it, instead of a copy
// of 0 per value.
__device__ __forceinline__ void set_pair(unsigned long long& v, unsigned lo, unsigned hi,
                                         bool hi_zero) {
  if (hi_zero)
    v = (v & 0xffffffff00000000ull) | lo;
  else
    v = ((unsigned long long)hi << 32) | lo;
}

// Pass 2 over one row group, in the groups of pass 1 (ng groups of four
// ending at the row's last pixel): the first group's pixels before the row
// are skipped by a mask per pixel, the others go under one mask per group.
// lw = log2 of the event multiplier in fixed-point units.  FOLD: the
// multiplier rides in the chain's kernel values; otherwise (MLEM events whose
// folded values could overflow a float) x = (K rho / lambda) 2^bits is formed
// per pixel from a direct exponential, without the large multiplier.  SPLIT:
// values reach 2^62 and carry a further 2^-32 until converted.  RPI (with
// FOLD and without SPLIT): values below 2^31, rounded by the conversion
// itself.  The conversion is inline assembly, which the compiler does not
// separate from a transcendental result by the wait state gfx950 needs; the
// values it reads come from a multiplication by the table.
template <int HP, bool RHO, bool FOLD, bool SPLIT, bool SUB, bool RPI = false,
          typename AccT = unsigned long long>
__device__ __forceinline__ void backproject_rows(const RowSeg& s, float du, float lw,
                                                 float lam, float unit_scale, ConstF* gtab,
                                                 const float* rho0, AccT* acc_w) {
  static_assert(!RPI || (FOLD && !SPLIT), "the rounding conversion reads chain values");
  constexpr bool k32 = sizeof(AccT) == 4;
  static_assert(!k32 || (RPI && !SUB), "32-bit windows take single rounded values");
  const int ng = (s.len + kUnroll - 1) >> 2;
  const int pad = ng * kUnroll - s.len;
  const float u0 = fmaf(-(float)pad, du, s.u0);
  const int w0 = s.w0 - pad * HP;
  const float* rp = s.len > 0 ? rho0 + w0 : rho0 - kSpareCols * HP;
  const int lastg = max(ng - 1, 0);
  const int maxng = s.mg;
  const float e_add = SPLIT ? lw - 32.0f : lw;
  const float uo[kUnroll] = {u0, u0 + du, fmaf(2.0f, du, u0), fmaf(3.0f, du, u0)};
  unsigned lo[kUnroll], hi[kUnroll];
  unsigned long long pv[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    pv[u] = 0ull;
    asm volatile("" : "+v"(pv[u]));
  }
  Chain c = chain_start(u0, s.lsc, s.b1, e_add);
  auto group = [&](int gk, bool first) {
    float rv[kUnroll] = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (RHO) read4<HP>(rp + min(gk, lastg) * (kUnroll * HP), rv);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      float x;
      if constexpr (FOLD) {
        x = fixed_arg<RHO, SPLIT, RPI>(c.p[u] * gtab[gk * kUnroll + u], rv[u]);
        c.p[u] *= c.b4;
      } else {
        // unit_scale = 2^bits (2^(bits - 32) when SPLIT)
        const float uu = fmaf((float)(gk * kUnroll), du, uo[u]);
        const float kr = __builtin_amdgcn_exp2f(fmaf(-uu, uu, s.lsc)) * rv[u];
        x = SPLIT ? (kr / lam) * unit_scale : fmaf(kr / lam, unit_scale, 0.5f);
      }
      to_fixed<SPLIT, RPI>(x, lo[u], hi[u]);
      set_pair(pv[u], lo[u], hi[u], !SPLIT);
    }
    if (gk < ng) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        if (!first || u == kUnroll - 1 || u >= pad) {
          if constexpr (k32)
            lds_accumulate32(acc_w + (w0 + (gk * kUnroll + u) * HP), lo[u]);
          else
            lds_accumulate<SUB>(acc_w + (w0 + (gk * kUnroll + u) * HP), pv[u]);
        }
    }
  };
  if (maxng > 0) group(0, true);
  for (int gk = 1; gk < maxng; ++gk) group(gk, false);
}

// Pass 1 of the analytic kernel over one row group, in the groups of
// forward_rows (ng groups of four ending at the row's last pixel): s0 +=
// sum K rho (rho: the window's values), kmax = the row's largest K.  Pixels
// before the row or past its end get K = 0 by a select (their K may not be
// finite).
template <int HP>
__device__ __forceinline__ void forward_rows_analytic(const RowSeg& s, const RowA& ra, float pw,
                                                      float c, float bq, const float* rho0,
                                                      float& s0, float& kmax) {
  const int ng = (s.len + kUnroll - 1) >> 2;
  const int pad = ng * kUnroll - s.len;
  const float* rp = s.len > 0 ? rho0 + (s.w0 - pad * HP) : rho0 - kSpareCols * HP;
  const int lastg = max(ng - 1, 0);
  const int maxng = s.mg;
  float sum = 0.0f, m = 0.0f;
  for (int gk = 0; gk < maxng; ++gk) {
    float rv[kUnroll];
    read4<HP>(rp + min(gk, lastg) * (kUnroll * HP), rv);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = gk * kUnroll + u - pad;  // the pixel's index in the row
      const float kv = analytic_at(c, bq, ra, fmaf((float)k, pw, ra.pq0));
      const float kk = (gk < ng && k >= 0) ? kv : 0.0f;
      sum = fmaf(kk, rv[u], sum);
      m = fmaxf(m, kk);
    }
  }
  s0 += sum;
  kmax = fmaxf(kmax, m);
}

// Pass 2 of the analytic kernel over one row group: x = K 2^lw (the event
// multiplier in fixed-point units, applied as two factors so that neither
// overflows), rounded by the conversion itself (values below 2^31) or, with
// SPLIT, carried with a further 2^-32 and converted in two halves (below
// 2^62); the same masks as backproject_rows.
template <int HP, bool SPLIT, typename AccT = unsigned long long>
__device__ __forceinline__ void backproject_rows_analytic(const RowSeg& s, const RowA& ra,
                                                          float pw, float c, float bq,
                                                          float lw, AccT* acc_w) {
  constexpr bool k32 = sizeof(AccT) == 4;
  static_assert(!k32 || !SPLIT, "32-bit windows take single rounded values");
  const int ng = (s.len + kUnroll - 1) >> 2;
  const int pad = ng * kUnroll - s.len;
  const int w0 = s.w0 - pad * HP;
  const int maxng = s.mg;
  const float e_add = SPLIT ? lw - 32.0f : lw;
  const float f1 = __builtin_amdgcn_exp2f(0.5f * e_add);
  const float f2 = __builtin_amdgcn_exp2f(e_add - 0.5f * e_add);
  unsigned lo[kUnroll], hi[kUnroll];
  unsigned long long pv[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    pv[u] = 0ull;
    asm volatile("" : "+v"(pv[u]));
  }
  for (int gk = 0; gk < maxng; ++gk) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = gk * kUnroll + u - pad;
      const float kv = analytic_at(c, bq, ra, fmaf((float)k, pw, ra.pq0));
      const float kk = (gk < ng && k >= 0) ? kv : 0.0f;
      const float x = (kk * f1) * f2;
      to_fixed<SPLIT, !SPLIT>(x, lo[u], hi[u]);
      set_pair(pv[u], lo[u], hi[u], !SPLIT);
    }
    if (gk < ng) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        if (gk > 0 || u >= pad) {
          if constexpr (k32)
            lds_accumulate32(acc_w + (w0 + (gk * kUnroll + u) * HP), lo[u]);
          else
            lds_accumulate<false>(acc_w + (w0 + (gk * kUnroll + u) * HP), pv[u]);
        }
    }
  }
}

}  // namespace

// A row-table entry, read once per iteration: a streaming (non-temporal)
// load, so that the table's tens of GB per launch do not evict the image
// lines the window flushes add to.
__device__ __forceinline__ RowEntry load_entry(const RowEntry* p) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  const v4u v = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(p));
  return RowEntry{v.x, __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w)};
}

// One workgroup of kWaves waves per work item (a chunk of events sharing a
// spatial bin).  The LDS windows are column-major with column stride HP, a
// multiple of 64 at least the window height: the 64 lanes of a wave walk 64
// different rows, so pixel (row, col) sits in bank (row + const) mod 64
// whatever column each lane has reached, and every LDS access of the walks
// is free of bank conflicts.
//
// BAND = 1, 2: the banded form for windows too tall for LDS.  Work items then
// cover one band of kBandRows image rows (and, for windows too wide as well,
// one band of 2^kColBandShift columns), and an event is walked in every band
// its box reaches, its rows clipped to the band's columns.  BAND = 1 runs pass 1 and stores the band's partial
// (lambda, sigma_z and sigma_dl sums, largest row factor) per event and
// band; k_band_combine adds them in a fixed order; BAND = 2 runs pass 2
// with the complete lambda.
//
// PACK = 2, 4 (work items whose boxes are at most 32 or 16 rows tall): a wave
// takes PACK events at a time, one per aligned segment of 64 / PACK lanes, so
// that short events do not leave most lanes idle; the denominator is then a
// segment sum and every per-event value lives per lane.
//
// LEAN: the walks under the default options, fixed at compile time: pass 1
// by Horner's rule where the mode has it, rows from their roots (no scan)
// and, unbanded, from the row table with its kernel values.  Without the
// other options' code the kernel keeps its uniform values in SGPRs instead
// of spilling them to VGPR lanes (the 64-row-stride kernels have 78 SGPRs at
// 8 waves per SIMD), and the 128-row stride kernel of the iteration drops
// from 95 to 54 VGPRs.  The 128-row LEAN kernels of the 32-bit-window
// iterations are compiled for 8 waves per SIMD (at most 64 VGPRs: the
// analytic one takes 61 instead of 67), so that two workgroups share a CU
// wherever the windows leave the LDS for it.
template <int MODE, bool COMP, int HP, int BAND, int PACK, bool LEAN>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(
    PACK > 1 || HP == 64 || (LEAN && HP == 128 && (MODE == MODE_MLEM_K || MODE == MODE_MLEM_KA))
        ? 8
        : 1)))
void k_strip_fused(FusedArgs a) {
  static_assert(PACK == 1 || (BAND == 0 && (MODE == MODE_MLEM || MODE == MODE_MLEM_K ||
                                            MODE == MODE_MLEM_KA || MODE == MODE_GRAD)),
                "packed events: the unbanded iteration and gradient");
  static_assert(MODE != MODE_MLEM_KA || BAND == 0, "the analytic walks are unbanded");
  constexpr bool kAnalytic = MODE == MODE_MLEM_KA;
  constexpr int kSeg = 64 / PACK;  // lanes per event
  // row groups kept in registers between the passes: every group of an
  // event (a box has at most HP rows, a band's rows at most HP = kBandRows)
  constexpr int kCached = HP / 64;
  constexpr bool kPass1 = BAND == 1 || (BAND == 0 && MODE != MODE_BACKWARD);
  constexpr bool kPass2 = BAND == 2 || (BAND == 0 && MODE != MODE_FORWARD);
  constexpr bool kRho = kPass1 || (kPass2 && MODE == MODE_MLEM);  // rho window in LDS
  constexpr bool kSigma = kPass1 && (MODE == MODE_FORWARD || MODE == MODE_GRAD);
  // pass-2 values K / lambda_e, without rho
  constexpr bool kNoRho2 = MODE == MODE_GRAD || MODE == MODE_MLEM_K || MODE == MODE_MLEM_KA;
  constexpr int kThreads = 64 * kWaves;
  // pass 1's group loop unrolled four times, except with the sigma sums in
  // the two-pass kernel, where the extra registers slowed the whole kernel
  constexpr int kPass1Unroll = (kSigma && BAND == 0) ? 1 : 4;
  // pass 1 by Horner's rule (forward_rows_h): the Gaussian walks with plain
  // sums and no sigma moments (the chain issues fewer instructions for those);
  // a.horner = 0 keeps the chain (forward_rows) for comparison
  constexpr bool kHorner = kPass1 && !COMP && !kAnalytic && !kSigma;
  const bool horner = kHorner && (LEAN || a.horner != 0);
  using U = unsigned long long;
  // the iteration's K / lambda sums accumulate in 32-bit windows (the item's
  // unit and cap below keep every window sum below 2^32); the gradient's (its
  // image term sum K / lambda - 1 cancels near the MLEM fixed point, so it
  // keeps the finer 64-bit unit), the K rho / lambda and the signed sums keep
  // 64 bits
  constexpr bool kAcc32 = MODE == MODE_MLEM_K || MODE == MODE_MLEM_KA;
  using AccT = std::conditional_t<kAcc32, uint32_t, U>;
  extern __shared__ __align__(16) unsigned char smem[];
  const StripConst& sg = a.sg;
  const StripConst& sk = a.sk;
  const WorkItem it = a.items[blockIdx.x];
  const int ww = it.ww, wh = it.wh;
  const int npx = ww * HP;
  // the rho window has four spare zero columns in front: a walk's first
  // group starts up to three pixels before its row, and an empty row reads
  // the four of them (values discarded); rho0 is the window's column 0
  const int nrho = (ww + kSpareCols) * HP;
  // the first kLdsHead bytes hold two floats per wave (the window sums below)
  float* wred = reinterpret_cast<float*>(smem);
  float* rho_w = reinterpret_cast<float*>(smem + kLdsHead);
  const float* rho0 = rho_w + kSpareCols * HP;
  AccT* acc_w = reinterpret_cast<AccT*>(
      smem + kLdsHead + (kRho ? (((size_t)nrho * 4 + 15) & ~(size_t)15) : 0));
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = __lane_id();
  const int seg = PACK > 1 ? lane / kSeg : 0;      // the lane's event of the wave's PACK
  const int lane_s = PACK > 1 ? lane - seg * kSeg : lane;  // the lane's row in it

  // Window load and flush visit tiles of 16 columns x 4 rows per wave, so
  // that the global reads and atomics of a wave cover 64-byte row segments.
  const int tc_lane = lane & 15, tr_lane = lane >> 4;
  // the thread's shares of the window's sums of rho and rho^2 (kAcc32)
  float vsum = 0.0f, vsq = 0.0f;
  if constexpr (kRho) {
    // tile columns from -16: tile 0 holds the spare columns
    const int ntc = ((ww + 15) >> 4) + 1, ntr = HP / 4;
    for (int t = wave; t < ntc * ntr; t += kWaves) {
      const int tr = t / ntc;
      const int c = (t - tr * ntc) * 16 + tc_lane - 16, r = tr * 4 + tr_lane;
      if (c < -kSpareCols || c >= ww) continue;
      float v = 0.0f;
      if (r < wh && c >= 0) {
        const size_t p = (size_t)(it.wy0 + r) * sg.W + (it.wx0 + c);
        v = a.rho[p];
        if (kPass1 && a.rho_weight) v *= a.rho_weight[p];  // the variant's rho s
        if constexpr (kAcc32) {
          vsum += v;
          vsq = fmaf(v, v, vsq);
        }
      }
      rho_w[(c + kSpareCols) * HP + r] = v;
    }
  } else if constexpr (kAcc32) {
    // pass 2 alone (banded): the window's rho for its sum only
    for (int i = threadIdx.x; i < ww * wh; i += kThreads) {
      const int r = i / ww, c = i - r * ww;
      const float v = a.rho[(size_t)(it.wy0 + r) * sg.W + (it.wx0 + c)];
      vsum += v;
      vsq = fmaf(v, v, vsq);
    }
  }
  if constexpr (kPass2) {
    for (int i = threadIdx.x; i < npx; i += kThreads) acc_w[i] = 0;
  }
  if constexpr (kAcc32) {
    vsum = wave_sum(vsum);
    vsq = wave_sum(vsq);
    if (lane == 0) {
      wred[wave] = vsum;
      wred[kWaves + wave] = vsq;
    }
  }
  __syncthreads();
  // The work item's fixed-point unit 2^-k of the K / lambda sums (kAcc32).
  // Its n events add to a window pixel at most once each, so values below
  // 2^vcap = 2^kLog2Sum32 / n keep every window sum below 2^32; rows whose
  // values could exceed it add to the image directly.  A typical value,
  // 1 / (rho x ellipse pixels), sits kItemHeadroomBits below the cap, with
  // rho the window's rho-weighted mean sum rho^2 / sum rho: the level an
  // event's denominator sees, which empty pixels do not dilute.  The window
  // follows the local activity, so cold and hot regions each get their own
  // resolution.  An empty window keeps the launch's unit.
  // (and below 2^30.5, the range of the rounding conversion to int32)
  const float vcap =
      fminf(kLog2Sum32 - __builtin_amdgcn_logf((float)max(it.end - it.begin, 1)), 30.5f);
  float item_bits = a.log2_scale;
  if constexpr (kAcc32) {
    float wsum = 0.0f, wsq = 0.0f;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      wsum += wred[w];
      wsq += wred[kWaves + w];
    }
    const float mean = wsum > 0.0f ? wsq / wsum : 0.0f;
    if (mean > 0.0f)
      item_bits = fminf(fmaxf(floorf(vcap - kItemHeadroomBits + __builtin_amdgcn_logf(mean) +
                                     a.log2_ell),
                              -60.0f),
                        60.0f);
  }
  const double item_unit = kAcc32 ? (double)__builtin_amdgcn_exp2f(-item_bits)
                                  : a.inv_fixed_scale;

  const float s_pix = a.s_pix;  // sqrt(isz log2 e)
  const bool scan = !LEAN && a.row_scan != 0;
  // The row table (unbanded walks): a lane's entry for row group gg of
  // sorted event e is e * stride + gg * 64 + its lane (its row in the event
  // when packed), in the order the walks visit the rows.  The entries of
  // groups 0 and 1 (when the stride has them) of the wave's next events load
  // while the current ones are walked.
  const bool use_tab = BAND == 0 && (LEAN || a.rowtab != nullptr);
  const bool tab_kernel = LEAN || a.rowtab_kernel != 0;
  constexpr int kPre = PACK == 1 && kCached > 1 ? 2 : 1;
  const bool pre1 = kPre > 1 && a.rowtab_stride > 64;
  RowEntry pre[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) pre[k] = RowEntry{0u, 0.0f, 0.0f, 0.0f};
  // a lane's entries sit at the wave step's first event (a uniform base)
  // plus loff; the table holds PACK - 1 spare events past the last, for the
  // idle segments of a step at the end of the events
  const uint32_t loff = (uint32_t)(seg * (int)a.rowtab_stride + lane_s);
  auto prefetch = [&](int e0n) {
    const RowEntry* wb = a.rowtab + (size_t)e0n * a.rowtab_stride;
    pre[0] = load_entry(wb + loff);
    if (pre1) pre[kPre - 1] = load_entry(wb + loff + 64u);
  };
  if (use_tab && it.begin + wave * PACK < it.end) prefetch(it.begin + wave * PACK);
  const float du = a.du;        // u step per column, pw s_pix
  ConstF* gtab = (ConstF*)a.gtab;  // uniform reads through the scalar cache
  ConstF* mtab = (ConstF*)a.mtab;
  const float inv_pw = 1.0f / sg.pw;
  const float inv_ph = 1.0f / sg.ph;
  const float cz_pix = 2.0f / (kLog2e * sk.sigma_z);  // 2 isz pq^2 / sz = cz u^2
  // 2^k, or 2^(k - 32) for the split conversion (MLEM's unfolded path)
  const float unit_scale = __builtin_amdgcn_exp2f(a.log2_scale - (a.split ? 32.0f : 0.0f));
  double ll = 0.0, gsz = 0.0, gsdl = 0.0, mmax = 0.0, msum = 0.0;

  for (int e0 = it.begin + wave * PACK; e0 < it.end; e0 += kWaves * PACK) {
    const int e = e0 + seg;
    const bool live = PACK == 1 || e < it.end;  // lanes past the item's last event idle
    const EventGeom g = a.sgeom[live ? e : e0];
    // box rows walked here: all of them, or those in this item's band
    // and columns: the box's, or those in this item's column band
    int ry0 = g.y0, ry1 = live ? g.y1 : g.y0, cx0 = g.x0, cx1 = g.x1;
    if constexpr (BAND != 0) {
      ry0 = max(g.y0, it.wy0);
      ry1 = min(g.y1, it.wy0 + wh);
      cx0 = max(g.x0, it.wx0);
      cx1 = min(g.x1, it.wx0 + ww);
      if (ry1 <= ry0 || cx1 <= cx0) continue;
    }
    const int nrows = ry1 - ry0;
    // packed events have at most kSeg rows: one row group for the wave
    const int ngroups = PACK > 1 ? 1 : (nrows + 63) >> 6;
    const int row0 = (ry0 - it.wy0) - it.wx0 * HP;  // window index of (ry0, 0)
    const int orig = a.sidx[live ? e : e0];
    // this event's entries of groups 0 and 1 (loaded with the last one);
    // those of the next event start loading now
    RowEntry cur[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) cur[k] = pre[k];
    if (use_tab && e0 + kWaves * PACK < it.end) prefetch(e0 + kWaves * PACK);
    // without the table: the per-event constants of the boundary search
    EventRows ev{0.0f, 0.0f, 0.0f, 0.0f};
    if (!use_tab) ev = event_rows(sg, g, inv_pw);
    // the lane's entry of row group gg (gi: its unrolled index)
    auto entry = [&](int gi, int gg) -> RowEntry {
      if (!use_tab) return RowEntry{0u, 0.0f, 0.0f, 0.0f};
      if (gi < kPre && (gi == 0 || pre1)) return cur[gi < kPre ? gi : 0];
      return load_entry(a.rowtab + (size_t)e0 * a.rowtab_stride + loff + (uint32_t)(gg * 64));
    };
    const int m0 = group0_first_row(sg, g, ry0, nrows, inv_ph);
    RowSeg cache[kCached];
    RowA acache[kAnalytic ? kCached : 1];

    // ---- pass 1: denominator (and sigma sums) -------------------------------
    float lam = 0.0f;
    if constexpr (kPass1) {
      float acc = 0.0f, comp = 0.0f, sz = 0.0f, sdl = 0.0f, lmax = -1e30f;
#pragma unroll
      for (int gi = 0; gi < kCached; ++gi) {
        for (int gg = gi; gg < gi + 1; ++gg) {
          if (gg >= ngroups) break;
          const int j = gg * 64 + lane - 64;
          const int r = PACK > 1 ? lane_s : gg == 0 ? m0 + lane : (j < m0 ? j : j + 64);
          RowSeg s{0, 0, 0.0f, -1e30f, 1.0f, 0.0f, 0.0f};
          RowA ra{0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
          if ((PACK > 1 ? lane_s : gg * 64 + lane) < nrows)
            s = row_setup<kSigma, kAnalytic>(sg, sk, g, ev, ry0 + r, row0 + r, HP, s_pix, du,
                                             scan, cx0, cx1, &ra, use_tab, entry(gi, gg),
                                             tab_kernel);
          set_walk_groups(s);
          float s0 = 0.0f, s1 = 0.0f;
          if constexpr (kAnalytic) {
            // the row's largest K bounds its pass-2 values (the chain's row
            // factor lsc does so for the Gaussian)
            float km = 0.0f;
            forward_rows_analytic<HP>(s, ra, sg.pw, sk.c, 2.0f * sk.isz, rho0, s0, km);
            s.lsc = km > 0.0f ? __builtin_amdgcn_logf(km) : -1e30f;
            acache[gi] = ra;
          } else {
            if (horner) {
              // the zero columns at the lane's window row
              const float* zp = rho_w + ((ry0 + r - it.wy0) & 63);
              forward_rows_h<HP, 2>(s, du, gtab, rho0, zp, s0);
            } else {
              forward_rows<HP, COMP, kSigma, kPass1Unroll>(s, du, gtab, mtab, rho0, s0, comp,
                                                           s1);
            }
          }
          cache[gi] = s;
          if constexpr (COMP) {
            float sum, err;
            two_sum(acc, s0, sum, err);
            acc = sum;
            comp += err;
          } else {
            acc += s0;
          }
          if constexpr (kSigma) {
            sz += fmaf(s.dz_row, s0, cz_pix * s1);
            sdl += s.ddl_row * s0;
            lmax = fmaxf(lmax, s.lsc);
          }
        }
      }
      lam = COMP ? seg_sum_comp<kSeg>(acc, comp) : seg_sum<PACK>(acc);
      if constexpr (BAND == 1) {
        // this band's partial sums of the event
        const float wsz = kSigma ? wave_sum(sz) : 0.0f;
        const float wsdl = kSigma ? wave_sum(sdl) : 0.0f;
        const float lmx = kSigma ? wave_max_f(lmax) : 0.0f;
        if (lane == 0) {
          const int slot = (it.pad - g.y0 / kBandRows) * a.ncb +
                           ((it.wx0 >> a.cshift) - (g.x0 >> a.cshift));
          float* p = a.part + ((size_t)e * a.nbands + slot) * 4;
          p[0] = lam;
          p[1] = wsz;
          p[2] = wsdl;
          p[3] = lmx;
        }
        continue;
      }
      if (lane_s == 0 && live) a.lambda[orig] = lam;
      if constexpr (kSigma) {
        // weight of this event: upstream g_e or 1/lambda_e (log-likelihood)
        const float ge =
            a.g ? a.g[orig] : (lam > 0.0f ? __builtin_amdgcn_rcpf(lam) : 0.0f);
        // every lane adds g_e times its rows' sigma sums in doubles; the
        // lanes are summed once, at the end, instead of two wave sums here
        gsz = fma((double)ge, (double)sz, gsz);
        gsdl = fma((double)ge, (double)sdl, gsdl);
        float kmax = 0.0f;
        if constexpr (MODE == MODE_FORWARD) kmax = __builtin_amdgcn_exp2f(wave_max_f(lmax));
        if (lane_s == 0 && live) {
          // ln lambda from the hardware log2; the sum over events is in doubles
          if (lam > 0.0f) ll += (double)__builtin_amdgcn_logf(lam) * 0.69314718055994531;
          if constexpr (MODE == MODE_FORWARD) {
            const double m = (double)fabsf(ge) * (double)kmax;
            mmax = fmax(mmax, m);
            msum += m;
          }
        }
      }
    }

    // ---- pass 2: backprojection into the LDS window --------------------------
    if constexpr (kPass2) {
      if constexpr (BAND == 2) lam = a.lam_sorted[e];
      // lw = log2 of the event multiplier in fixed-point units
      float lw;
      bool neg = false;
      if constexpr (MODE == MODE_MLEM || kNoRho2) {
        // packed: the lanes of an event without a positive denominator walk
        // empty rows (below)
        if (PACK == 1 ? !(lam > 0.0f) : !__any(lam > 0.0f)) continue;
        lw = log2_shift(item_bits, lam > 0.0f ? lam : 1.0f, -1.0f);
      } else if (a.g) {
        const float ge = a.g[orig];
        if (ge == 0.0f) continue;
        neg = ge < 0.0f;  // uniform: one sign per event
        lw = log2_shift(a.log2_scale, fabsf(ge), 1.0f);
      } else {
        const float l = a.lambda[orig];
        if (!(l > 0.0f)) continue;
        lw = log2_shift(a.log2_scale, l, -1.0f);
      }
#pragma unroll
      for (int gi = 0; gi < kCached; ++gi) {
        for (int gg = gi; gg < gi + 1; ++gg) {
          if (gg >= ngroups) break;
          RowSeg s;
          RowA ra{0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
          const int j = gg * 64 + lane - 64;
          const int r = PACK > 1 ? lane_s : gg == 0 ? m0 + lane : (j < m0 ? j : j + 64);
          if constexpr (kPass1) {
            s = cache[gi];
            if constexpr (kAnalytic) ra = acache[gi];
            if (PACK > 1 && !(lam > 0.0f)) s.len = 0;
          } else {
            s = RowSeg{0, 0, 0.0f, -1e30f, 1.0f, 0.0f, 0.0f};
            if ((PACK > 1 ? lane_s : gg * 64 + lane) < nrows && (PACK == 1 || lam > 0.0f))
              s = row_setup<false>(sg, sk, g, ev, ry0 + r, row0 + r, HP, s_pix, du, scan, cx0,
                                   cx1, nullptr, use_tab, entry(gi, gg), tab_kernel);
            set_walk_groups(s);
          }
          if constexpr (MODE == MODE_MLEM) {
            // K rho 2^bits / lambda <= 2^bits, but the chain's K 2^bits /
            // lambda alone (up to 2^(lsc + lw + 20) along a row) overflows a
            // float when lambda is tiny
            const bool fold = !__any(s.lsc + lw > 100.0f);
            if (a.split) {
              // 2^-40 units (the deterministic accumulation): values up to 2^40
              if (fold)
                backproject_rows<HP, true, true, true, false>(s, du, lw, lam, unit_scale,
                                                              gtab, rho0, acc_w);
              else
                backproject_rows<HP, true, false, true, false>(s, du, lw, lam, unit_scale,
                                                               gtab, rho0, acc_w);
            } else if (fold) {
              backproject_rows<HP, true, true, false, false>(s, du, lw, lam, unit_scale,
                                                             gtab, rho0, acc_w);
            } else {
              backproject_rows<HP, true, false, false, false>(s, du, lw, lam, unit_scale,
                                                              gtab, rho0, acc_w);
            }
          } else if constexpr (kAnalytic) {
            const float vmax = s.lsc + lw;
            if (!__any(vmax > vcap)) {
              backproject_rows_analytic<HP, false>(s, ra, sg.pw, sk.c, 2.0f * sk.isz, lw,
                                                   acc_w);
            } else if (s.len > 0) {
              const int col0 = it.wx0 + (s.w0 - (row0 + r + it.wx0 * HP)) / HP;
              const size_t p0 = (size_t)(ry0 + r) * sg.W + col0;
              for (int k = 0; k < s.len; ++k) {
                const float kv = analytic_at(sk.c, 2.0f * sk.isz, ra,
                                             fmaf((float)k, sg.pw, ra.pq0));
                if (a.acc_kind == GLOBAL_F64)
                  unsafeAtomicAdd(reinterpret_cast<double*>(a.acc) + p0 + k,
                                  (double)kv / (double)lam);
                else
                  unsafeAtomicAdd(reinterpret_cast<float*>(a.acc) + p0 + k,
                                  (float)((double)kv / (double)lam));
              }
            }
          } else if constexpr (kNoRho2) {
            // K / lambda has no bound: rows whose values stay below 2^vcap
            // fixed-point units (32-bit windows; 2^30.5 or, split, 2^52 in
            // the 64-bit ones) take the single rounding conversion, the rare
            // rest is added to the global image in its own format
            const float vmax = s.lsc + lw;
            if (!__any(vmax > (kAcc32 ? vcap : 30.5f))) {
              backproject_rows<HP, false, true, false, false, true>(
                  s, du, lw, 0.0f, unit_scale, gtab, rho0, acc_w);
            } else if (!kAcc32 && !__any(vmax > 52.0f)) {
              if constexpr (!kAcc32)
                backproject_rows<HP, false, true, true, false>(s, du, lw, 0.0f, unit_scale,
                                                               gtab, rho0, acc_w);
            } else if (s.len > 0) {
              const int col0 = it.wx0 + (s.w0 - (row0 + r + it.wx0 * HP)) / HP;
              const size_t p0 = (size_t)(ry0 + r) * sg.W + col0;
              for (int k = 0; k < s.len; ++k) {
                const float uu = fmaf((float)k, du, s.u0);
                const float kv = __builtin_amdgcn_exp2f(fmaf(-uu, uu, s.lsc));
                if (a.acc_kind == GLOBAL_F64)
                  unsafeAtomicAdd(reinterpret_cast<double*>(a.acc) + p0 + k,
                                  (double)kv / (double)lam);
                else
                  unsafeAtomicAdd(reinterpret_cast<float*>(a.acc) + p0 + k,
                                  (float)((double)kv / (double)lam));
              }
            }
          } else if (a.split) {
            // |g| K 2^k <= 2^62 / n_events by the choice of k
            if (neg)
              backproject_rows<HP, false, true, true, true>(s, du, lw, 0.0f, unit_scale,
                                                            gtab, rho0, acc_w);
            else
              backproject_rows<HP, false, true, true, false>(s, du, lw, 0.0f, unit_scale,
                                                             gtab, rho0, acc_w);
          } else {
            // |g| K 2^k <= 2^31 by the choice of k
            if (neg)
              backproject_rows<HP, false, true, false, true>(s, du, lw, 0.0f, unit_scale,
                                                             gtab, rho0, acc_w);
            else
              backproject_rows<HP, false, true, false, false>(s, du, lw, 0.0f, unit_scale,
                                                              gtab, rho0, acc_w);
          }
        }
      }
    }
  }

  if constexpr (kPass2) {
    __syncthreads();
    const int ntc = (ww + 15) >> 4, ntr = (wh + 3) >> 2;
    for (int t = wave; t < ntc * ntr; t += kWaves) {
      const int tr = t / ntc;
      const int c = (t - tr * ntc) * 16 + tc_lane, r = tr * 4 + tr_lane;
      if (c >= ww || r >= wh) continue;
      const AccT v = acc_w[c * HP + r];
      if (v == 0) continue;
      size_t p = (size_t)(it.wy0 + r) * sg.W + (it.wx0 + c);
      // 64-bit sums are signed (BACKWARD subtracts), 32-bit ones are not
      const double dv = kAcc32 ? (double)v : (double)(long long)v;
      if (a.acc_kind == GLOBAL_U64)
        atomicAdd(reinterpret_cast<U*>(a.acc) + p, (U)v);
      else if (a.acc_kind == GLOBAL_F64)
        unsafeAtomicAdd(reinterpret_cast<double*>(a.acc) + p, dv * item_unit);
      else
        unsafeAtomicAdd(reinterpret_cast<float*>(a.acc) + p, (float)(dv * item_unit));
    }
  }

  if constexpr (kSigma && BAND == 0) {
    // Lane 0 of each wave holds the wave's scalars (the sigma sums once its
    // lanes are added).  The workgroup's sums go
    // to its own slot in a fixed wave order: five atomics per wave onto the
    // same five addresses queue at one L2 channel, and each wave's wait for
    // them held its workgroup for milliseconds on small footprints.
    gsz = wave_sum_d(gsz);
    gsdl = wave_sum_d(gsdl);
    if constexpr (PACK > 1) ll = wave_sum_d(ll);  // held by each event's first lane
    __syncthreads();  // every wave is done with the LDS windows
    double* red = reinterpret_cast<double*>(smem);
    if (lane == 0) {
      red[wave * 5 + 0] = ll;
      red[wave * 5 + 1] = gsz;
      red[wave * 5 + 2] = gsdl;
      red[wave * 5 + 3] = mmax;
      red[wave * 5 + 4] = msum;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
      const int i = threadIdx.x;
      double s = red[i];
      for (int w = 1; w < kWaves; ++w) s = i == 3 ? fmax(s, red[w * 5 + i]) : s + red[w * 5 + i];
      a.scal_part[(size_t)blockIdx.x * kScalStride + i] = s;
    }
  }
}

// Adds the per-workgroup partials of the scalars (sum log lambda, the two
// sigma sums, max and sum of |g| K; kScalStride doubles apart) to scalars,
// in a fixed order: thread t sums blocks t, t + 256, ..., then a fixed tree.
__global__ __launch_bounds__(256) void k_scalars_finish(const double* part, int64_t nblocks,
                                                        double* scalars) {
  __shared__ double red[5][256];
  const int t = threadIdx.x;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t b = t; b < nblocks; b += 256) {
    const double* p = part + b * kScalStride;
    v[0] += p[0];
    v[1] += p[1];
    v[2] += p[2];
    v[3] = fmax(v[3], p[3]);
    v[4] += p[4];
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) red[i][t] = v[i];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int i = 0; i < 5; ++i)
        red[i][t] = i == 3 ? fmax(red[i][t], red[i][t + s]) : red[i][t] + red[i][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    scalars[0] += red[0][0];
    scalars[1] += red[1][0];
    scalars[2] += red[2][0];
    scalars[3] = fmax(scalars[3], red[3][0]);
    scalars[4] += red[4][0];
  }
}

// Per-event pixel counts of the fused walks, one wave per event: n_box of
// the clamped box and n_ell, the sum of the row intervals row_setup finds
// (the pixels the reference's box scan accepts, reconstruction.h:236-247).
__global__ __launch_bounds__(256) void k_row_counts(StripConst sg, const EventGeom* geom,
                                                    int64_t n, int scan, int32_t* nbox,
                                                    int32_t* nell) {
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= n) return;  // uniform per wave
  const int lane = __lane_id();
  const EventGeom g = geom[e];
  const int nrows = max(g.y1 - g.y0, 0), ncols = max(g.x1 - g.x0, 0);
  const EventRows ev = event_rows(sg, g, 1.0f / sg.pw);
  int cnt = 0;
  if (ncols > 0) {
    for (int r = lane; r < nrows; r += 64)
      cnt += row_setup<false>(sg, sg, g, ev, g.y0 + r, 0, 0, 1.0f, 0.0f, scan != 0, g.x0, g.x1)
                 .len;
  }
  for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
  if (lane == 0) {
    nbox[e] = nrows * ncols;
    nell[e] = cnt;
  }
}

hipError_t launch_row_counts(hipStream_t st, const StripConst& sg, const EventGeom* geom,
                             int64_t n, int scan, int32_t* nbox, int32_t* nell) {
  // one wave per event: slices of 2^24 events keep the grid's work-item
  // count within 32 bits
  constexpr int64_t kSlice = (int64_t)1 << 24;
  for (int64_t e0 = 0; e0 < n; e0 += kSlice) {
    const int64_t m = std::min(kSlice, n - e0);
    hipLaunchKernelGGL(k_row_counts, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, sg,
                       geom + e0, m, scan, nbox + e0, nell + e0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// The row table, one wave per sorted event: the entry e * stride + slot
// holds the walk (columns, u0, row factor, chain ratio, as row_setup forms
// them) of the box row that the fused kernel's lane visits at that slot:
// slot gg * 64 + lane for row group gg (group 0 the 64 rows around the
// ellipse centre from group0_first_row, then the rows above and below it in
// order); slots past the box's rows hold empty entries.
__global__ __launch_bounds__(256) void k_row_table(StripConst sg, StripConst sk, float s_pix,
                                                   float du, const EventGeom* geom, int64_t n,
                                                   int scan, int64_t stride, RowEntry* tab) {
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= n) return;  // uniform per wave
  const int lane = __lane_id();
  const EventGeom g = geom[e];
  const int nrows = g.x1 > g.x0 ? max(g.y1 - g.y0, 0) : 0;
  const int m0 = group0_first_row(sg, g, g.y0, nrows, 1.0f / sg.ph);
  const EventRows ev = event_rows(sg, g, 1.0f / sg.pw);
  for (int slot = lane; slot < stride; slot += 64) {
    const int j = slot - 64;
    const int r = slot < 64 ? m0 + slot : (j < m0 ? j : j + 64);
    RowEntry t{0u, 0.0f, 0.0f, 0.0f};
    if (r < nrows) {
      // window index = column: w0 is the first ellipse column
      const RowSeg s =
          row_setup<false>(sg, sk, g, ev, g.y0 + r, 0, 1, s_pix, du, scan != 0, g.x0, g.x1);
      if (s.len > 0)
        t = RowEntry{((uint32_t)s.len << 16) | (uint32_t)(s.w0 - g.x0), s.u0, s.lsc, s.b1};
    }
    tab[e * stride + slot] = t;
  }
}

hipError_t launch_row_table(hipStream_t st, const StripConst& sg, const StripConst& sk,
                            float s_pix, float du, const EventGeom* sgeom, int64_t n, int scan,
                            int64_t stride, RowEntry* tab) {
  constexpr int64_t kSlice = (int64_t)1 << 24;
  for (int64_t e0 = 0; e0 < n; e0 += kSlice) {
    const int64_t m = std::min(kSlice, n - e0);
    hipLaunchKernelGGL(k_row_table, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, sg, sk,
                       s_pix, du, sgeom + e0, m, scan, stride, tab + e0 * stride);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------

size_t fused_lds_bytes(int max_ww, int col_stride, int mode, int band) {
  const size_t hp = (size_t)col_stride;
  const bool no_rho = mode == MODE_BACKWARD || (band == 2 && mode != MODE_MLEM);
  size_t rho = no_rho ? 0 : ((((size_t)max_ww + kSpareCols) * hp * 4 + 15) & ~(size_t)15);
  const bool acc32 = mode == MODE_MLEM_K || mode == MODE_MLEM_KA;
  size_t acc = mode == MODE_FORWARD ? 0 : (size_t)max_ww * hp * (acc32 ? 4 : 8);
  return kLdsHead + rho + acc + 16;
}

int fused_col_stride(int max_wh) {
  if (max_wh < 1) max_wh = 1;
  const int hp = (max_wh + 63) / 64 * 64;
  return hp <= kMaxColStride ? hp : 0;
}

namespace {

// Items [i_begin, i_end) with PACK events per wave at a time.
template <int MODE, bool COMP, int HP, int BAND, int PACK>
hipError_t launch_range(hipStream_t st, const FusedArgs& a, int i_begin, int i_end,
                        size_t lds) {
  if (i_end <= i_begin) return hipSuccess;
  // the default options' form (LEAN) when the launch has them (the row
  // table's rows were formed under the launch's row_scan)
  auto kern = k_strip_fused<MODE, COMP, HP, BAND, PACK, false>;
  const bool lean = a.horner != 0 && (BAND == 0 ? a.rowtab != nullptr && a.rowtab_kernel != 0
                                                : a.row_scan == 0);
  if (lean) kern = k_strip_fused<MODE, COMP, HP, BAND, PACK, true>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // 2^21 items of 1024 work-items per launch keep the grid within 32 bits
  constexpr int kSlice = 1 << 21;
  for (int i0 = i_begin; i0 < i_end; i0 += kSlice) {
    FusedArgs b = a;
    b.items = a.items + i0;
    b.scal_part = a.scal_part + (size_t)i0 * kScalStride;
    hipLaunchKernelGGL(kern, dim3((unsigned)std::min(kSlice, i_end - i0)),
                       dim3(64 * kWaves), lds, st, b);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

// The banded launches: every item in one range of one event per wave.
template <int MODE, bool COMP, int HP, int BAND>
hipError_t launch_t(hipStream_t st, const FusedArgs& a, int n_items, size_t lds) {
  return launch_range<MODE, COMP, HP, BAND, 1>(st, a, 0, n_items, lds);
}

// The unbanded launches, range by range (ItemRanges): the iteration and the
// gradient pack two or four of their events into a wave for the short boxes,
// and each range's LDS follows its widest window.
template <int MODE, bool COMP, int HP>
hipError_t launch_u(hipStream_t st, const FusedArgs& a, const ItemRanges& r) {
  constexpr bool kPackable = MODE == MODE_MLEM || MODE == MODE_MLEM_K ||
                             MODE == MODE_MLEM_KA || MODE == MODE_GRAD;
  for (int i = 0; i < r.n; ++i) {
    const int b = r.begin[i], e1 = std::min(r.end[i], r.n_items);
    const size_t lds = fused_lds_bytes(r.max_ww[i], HP, MODE);
    hipError_t e;
    if (kPackable && r.pack[i] == 4)
      e = launch_range<MODE, COMP, HP, 0, kPackable ? 4 : 1>(st, a, b, e1, lds);
    else if (kPackable && r.pack[i] == 2)
      e = launch_range<MODE, COMP, HP, 0, kPackable ? 2 : 1>(st, a, b, e1, lds);
    else
      e = launch_range<MODE, COMP, HP, 0, 1>(st, a, b, e1, lds);
    if (e != hipSuccess) return e;
  }
  if constexpr (MODE == MODE_FORWARD || MODE == MODE_GRAD) {
    hipLaunchKernelGGL(k_scalars_finish, dim3(1), dim3(256), 0, st, a.scal_part,
                       (int64_t)r.n_items, a.scalars);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <int MODE, bool COMP>
hipError_t launch_h(int col_stride, hipStream_t st, const FusedArgs& a, const ItemRanges& r) {
  switch (col_stride) {
    case 64: return launch_u<MODE, COMP, 64>(st, a, r);
    case 128: return launch_u<MODE, COMP, 128>(st, a, r);
    case 192: return launch_u<MODE, COMP, 192>(st, a, r);
    case 256: return launch_u<MODE, COMP, 256>(st, a, r);
    default: return hipErrorInvalidValue;
  }
}

// Sums each event's band partials in band order: lambda (sorted and original
// order) and, with sigma, the log-likelihood terms, the sigma derivative
// weighted by g_e (or 1/lambda_e) and the largest |g_e| K.  A fixed grid of
// kCombineBlocks blocks strides over the events; each block's sums go to its
// slot of scal_part (k_scalars_finish adds them up in a fixed order).
__global__ __launch_bounds__(256) void k_band_combine(const float* part, int64_t n,
                                                      int nbands, const int32_t* sidx,
                                                      const float* g, int sigma,
                                                      float* lam_sorted, float* lambda,
                                                      double* scal_part) {
  __shared__ double red[4][5];
  double ll = 0.0, gsz = 0.0, gsdl = 0.0, mmax = 0.0, msum = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    float lam = 0.0f, sz = 0.0f, sdl = 0.0f, lmx = -1e30f;
    const float* p = part + (size_t)e * nbands * 4;
    for (int j = 0; j < nbands; ++j) {
      lam += p[4 * j];
      sz += p[4 * j + 1];
      sdl += p[4 * j + 2];
      // slots of bands the event does not reach hold zeros
      if (p[4 * j] != 0.0f || p[4 * j + 3] != 0.0f) lmx = fmaxf(lmx, p[4 * j + 3]);
    }
    const int orig = sidx[e];
    lam_sorted[e] = lam;
    lambda[orig] = lam;
    if (sigma) {
      const float ge = g ? g[orig] : (lam > 0.0f ? 1.0f / lam : 0.0f);
      if (lam > 0.0f) ll += log((double)lam);
      gsz += (double)ge * sz;
      gsdl += (double)ge * sdl;
      const double m = (double)fabsf(ge) * (double)exp2f(lmx);
      mmax = fmax(mmax, m);
      msum += m;
    }
  }
  if (sigma) {
    for (int off = 32; off >= 1; off >>= 1) {
      ll += __shfl_xor(ll, off, 64);
      gsz += __shfl_xor(gsz, off, 64);
      gsdl += __shfl_xor(gsdl, off, 64);
      mmax = fmax(mmax, __shfl_xor(mmax, off, 64));
      msum += __shfl_xor(msum, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[w][0] = ll;
      red[w][1] = gsz;
      red[w][2] = gsdl;
      red[w][3] = mmax;
      red[w][4] = msum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < 4; ++i) {
        red[0][0] += red[i][0];
        red[0][1] += red[i][1];
        red[0][2] += red[i][2];
        red[0][3] = fmax(red[0][3], red[i][3]);
        red[0][4] += red[i][4];
      }
      double* p = scal_part + (size_t)blockIdx.x * kScalStride;
      for (int i = 0; i < 5; ++i) p[i] = red[0][i];
    }
  }
}

}  // namespace

hipError_t launch_strip_fused(hipStream_t st, int mode, int col_stride, int compensated,
                              const FusedArgs& a, const ItemRanges& r) {
  if (r.n_items <= 0) return hipSuccess;
  switch (mode) {
    case MODE_MLEM:
      return compensated ? launch_h<MODE_MLEM, true>(col_stride, st, a, r)
                         : launch_h<MODE_MLEM, false>(col_stride, st, a, r);
    case MODE_FORWARD:
      return compensated ? launch_h<MODE_FORWARD, true>(col_stride, st, a, r)
                         : launch_h<MODE_FORWARD, false>(col_stride, st, a, r);
    case MODE_GRAD: return launch_h<MODE_GRAD, false>(col_stride, st, a, r);
    case MODE_MLEM_K: return launch_h<MODE_MLEM_K, false>(col_stride, st, a, r);
    case MODE_MLEM_KA: return launch_h<MODE_MLEM_KA, false>(col_stride, st, a, r);
    default: return launch_h<MODE_BACKWARD, false>(col_stride, st, a, r);
  }
}

hipError_t launch_strip_banded(hipStream_t st, int mode, int band, const FusedArgs& a,
                               int n_items, size_t lds_bytes) {
  if (n_items <= 0) return hipSuccess;
  if (band == 1) {
    return mode == MODE_FORWARD
               ? launch_t<MODE_FORWARD, false, kBandRows, 1>(st, a, n_items, lds_bytes)
               : launch_t<MODE_MLEM, false, kBandRows, 1>(st, a, n_items, lds_bytes);
  }
  switch (mode) {
    case MODE_MLEM: return launch_t<MODE_MLEM, false, kBandRows, 2>(st, a, n_items, lds_bytes);
    case MODE_GRAD: return launch_t<MODE_GRAD, false, kBandRows, 2>(st, a, n_items, lds_bytes);
    case MODE_MLEM_K:
      return launch_t<MODE_MLEM_K, false, kBandRows, 2>(st, a, n_items, lds_bytes);
    default: return launch_t<MODE_BACKWARD, false, kBandRows, 2>(st, a, n_items, lds_bytes);
  }
}

hipError_t launch_band_combine(hipStream_t st, const FusedArgs& a, int64_t n, int sigma) {
  if (n <= 0) return hipSuccess;
  const int blocks = (int)std::min<int64_t>(kCombineBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(k_band_combine, dim3((unsigned)blocks), dim3(256), 0, st, a.part, n,
                     a.nbands, a.sidx, a.g, sigma, a.lam_sorted, a.lambda, a.scal_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !sigma) return e;
  hipLaunchKernelGGL(k_scalars_finish, dim3(1), dim3(256), 0, st, a.scal_part,
                     (int64_t)blocks, a.scalars);
  return hipGetLastError();
}

}  // namespace petstrip
