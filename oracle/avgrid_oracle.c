/*
 * avgrid_oracle.c -- FP64 CPU restatement of the reference standard gridder.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is linked by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nowhere else.  The product (astroviper_amd/) never calls it.
 *
 * Every function restates one reference function loop by loop, in the same
 * order and with the same rounding rules, in plain C99 compiled with
 * -O3 -ffp-contract=off (no fused multiply-add, no fast-math), so that the
 * FP64 arithmetic is the same IEEE sequence the reference's numba/NumPy code
 * performs.  Reference paths are relative to src/astroviper/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Schwab rational-approximation coefficients, M=6, alpha=1.
 * core/imaging/imaging_utils/gcf_prolate_spheroidal.py:116-127 */
static const double PS_P[2][5] = {
    {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
    {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
static const double PS_Q[2][3] = {{1.0000000e0, 8.212018e-1, 2.078043e-1},
                                  {1.0000000e0, 9.599102e-1, 2.918724e-1}};

/* prolate_spheroidal_function, gcf_prolate_spheroidal.py:102-158.
 * For each u: part/uend chosen by |u| (<0.75 -> 0/0.75, [0.75,1] -> 1/1.0,
 * otherwise the zero-initialised defaults part 0, uend 0), then
 * top = sum_k p[part,k] * pow(d,k), bot = sum_k q[part,k] * pow(d,k) with
 * d = u*u - uend*uend; grdsf = top/bot where bot > 0, forced to 0 for |u|>1.
 * np.power(float64, int) goes through libm pow(), as here.  Returns grdsf
 * and (1 - u*u) * grdsf. */
void ora_ps_function(const double *u_in, int64_t n, double *grdsf_out,
                     double *kern_out) {
  for (int64_t i = 0; i < n; ++i) {
    double u = fabs(u_in[i]);
    int part = 0;
    double uend = 0.0;
    if (u >= 0.0 && u < 0.75) {
      part = 0;
      uend = 0.75;
    }
    if (u >= 0.75 && u <= 1.0) {
      part = 1;
      uend = 1.0;
    }
    double delusq = u * u - uend * uend;
    double top = PS_P[part][0];
    for (int k = 1; k < 5; ++k) top += PS_P[part][k] * pow(delusq, (double)k);
    double bot = PS_Q[part][0];
    for (int k = 1; k < 3; ++k) bot += PS_Q[part][k] * pow(delusq, (double)k);
    double g = 0.0;
    if (bot > 0.0) g = top / bot;
    if (u > 1.0) g = 0.0;
    if (grdsf_out) grdsf_out[i] = g;
    if (kern_out) kern_out[i] = (1.0 - u * u) * g;
  }
}

/* create_prolate_spheroidal_kernel_1D, gcf_prolate_spheroidal.py:87-99.
 * out has ov*(sc+1) entries; entry i < ov*sc is the gridding kernel at
 * u = i / (sc*ov), the last ov entries are zero. */
void ora_ps_kernel_1d(int ov, int support, double *out) {
  int sc = support / 2;
  int n = ov * sc;
  int ntot = ov * (sc + 1);
  for (int i = 0; i < ntot; ++i) out[i] = 0.0;
  for (int i = 0; i < n; ++i) {
    double u = (double)i / (double)(sc * ov);
    ora_ps_function(&u, 1, NULL, &out[i]);
  }
}

/* Gridding-correction taper kernel_image of create_prolate_spheroidal_kernel,
 * gcf_prolate_spheroidal.py:72-79 with coordinates() at :161-163:
 * taper[i,j] = grdsf(|2*(i - nu//2)/nu|) * grdsf(|2*(j - nv//2)/nv|). */
void ora_ps_taper(int nu, int nv, double *out) {
  for (int i = 0; i < nu; ++i) {
    double cu = (double)(i - nu / 2) / (double)nu;
    double au = fabs(2.0 * cu), gu;
    ora_ps_function(&au, 1, &gu, NULL);
    for (int j = 0; j < nv; ++j) {
      double cv = (double)(j - nv / 2) / (double)nv;
      double av = fabs(2.0 * cv), gv;
      ora_ps_function(&av, 1, &gv, NULL);
      out[(int64_t)i * nv + j] = gu * gv;
    }
  }
}

/* standard_grid_jit, core/imaging/imaging_utils/standard_grid.py:246-432.
 *
 * grid      complex128 (interleaved re,im) [n_img_chan][n_img_pol][n_u][n_v],
 *           or float64 when complex_grid == 0 (do_psf only), accumulated.
 * sum_weight float64 [n_img_chan][n_img_pol], accumulated.
 * vis       complex128 [nt][nb][nchan][nvispol] (ignored when do_psf).
 * uvw       float64 [nt][nb][3];  weight float64 [nt][nb][nchan][npol].
 * chan_map  [nchan] image channel, pol_map [npol] image polarisation.
 * npol = len(pol_map) is the loop count over polarisations (:313).
 *
 * Loop order time -> baseline -> channel -> polarisation -> v-tap -> u-tap
 * (:318-404).  Centre index int(x + 0.5) truncates toward zero (:337-341),
 * offset index floor((ic - pos) * ov + 0.5) (:349-352), kernel index
 * |ov * tap + offset| (:384-396).  A sample is skipped if its weighted datum
 * is NaN or exactly zero (:378).  sum_weight gains weight * norm (:423-430).
 * do_imaging_weight additionally scatters to the conjugate position and
 * doubles the sum_weight contribution (:389-430); the conjugate taps are
 * bounds-checked here (the reference does not check them). */
void ora_standard_grid(double *grid, int complex_grid, double *sum_weight,
                       int do_psf, int do_imaging_weight, const double *vis,
                       int nvispol, const double *uvw, int nt, int nb,
                       const double *freq_chan, int nchan,
                       const int64_t *chan_map, const int64_t *pol_map,
                       int npol, const double *weight, int nwpol,
                       const double *cgk_1D, int n_img_chan, int n_img_pol,
                       int n_u, int n_v, double dl_u, double dl_v, int support,
                       int oversampling) {
  const double c = 299792458.0;
  const int support_center = support / 2;
  const int64_t uv_center_u = n_u / 2, uv_center_v = n_v / 2;
  const int start_support = -support_center;
  const int end_support = support - support_center;
  double *uv_scale_u = (double *)malloc(sizeof(double) * (size_t)(nchan + 1));
  double *uv_scale_v = (double *)malloc(sizeof(double) * (size_t)(nchan + 1));
  (void)n_img_chan;
  (void)n_img_pol;
  for (int ch = 0; ch < nchan; ++ch) {
    /* numpy: -(freq_chan * delta_lm[0] * n_uv[0]) / c  (:298-299) */
    uv_scale_u[ch] = -(freq_chan[ch] * dl_u * (double)n_u) / c;
    uv_scale_v[ch] = -(freq_chan[ch] * dl_v * (double)n_v) / c;
  }
  const int64_t plane = (int64_t)n_u * n_v;
  for (int t = 0; t < nt; ++t) {
    for (int b = 0; b < nb; ++b) {
      const int64_t row = (int64_t)t * nb + b;
      for (int ch = 0; ch < nchan; ++ch) {
        int64_t a_chan = chan_map[ch];
        double u = uvw[row * 3 + 0] * uv_scale_u[ch];
        double v = uvw[row * 3 + 1] * uv_scale_v[ch];
        if (isnan(u) || isnan(v)) continue;
        double u_pos = u + (double)uv_center_u;
        double v_pos = v + (double)uv_center_v;
        double u_pos_conj = -u + (double)uv_center_u;
        double v_pos_conj = -v + (double)uv_center_v;
        /* int(x + 0.5): truncation toward zero.  Values outside the int64
         * range cannot pass the bounds test below, so they are rejected
         * before the cast. */
        double xu = u_pos + 0.5, xv = v_pos + 0.5;
        if (!(xu > -9.0e18 && xu < 9.0e18 && xv > -9.0e18 && xv < 9.0e18))
          continue;
        int64_t u_ci = (int64_t)xu;
        int64_t v_ci = (int64_t)xv;
        int64_t u_ci_conj = (int64_t)(u_pos_conj + 0.5);
        int64_t v_ci_conj = (int64_t)(v_pos_conj + 0.5);
        if (!((u_ci + support_center < n_u) && (v_ci + support_center < n_v) &&
              (u_ci - support_center >= 0) && (v_ci - support_center >= 0)))
          continue;
        double u_offset = (double)u_ci - u_pos;
        int64_t u_coi = (int64_t)floor(u_offset * (double)oversampling + 0.5);
        double v_offset = (double)v_ci - v_pos;
        int64_t v_coi = (int64_t)floor(v_offset * (double)oversampling + 0.5);
        for (int ip = 0; ip < npol; ++ip) {
          const int64_t widx = ((row * nchan) + ch) * nwpol;
          double wd_re, wd_im, sel_weight;
          if (do_psf) {
            if (npol >= 2 && do_imaging_weight) {
              wd_re = (weight[widx + 0] + weight[widx + 1]) / 2.0;
              sel_weight = wd_re;
            } else {
              sel_weight = weight[widx + ip];
              wd_re = weight[widx + ip];
            }
            wd_im = 0.0;
          } else {
            sel_weight = weight[widx + ip];
            const int64_t vidx = (((row * nchan) + ch) * nvispol + ip) * 2;
            wd_re = vis[vidx] * weight[widx + ip];
            wd_im = vis[vidx + 1] * weight[widx + ip];
          }
          if (isnan(wd_re) || isnan(wd_im)) continue;
          if (wd_re == 0.0 && wd_im == 0.0) continue;
          int64_t a_pol = pol_map[ip];
          double *g = grid + ((a_chan * n_img_pol) + a_pol) * plane *
                                 (complex_grid ? 2 : 1);
          double norm = 0.0;
          for (int i_v = start_support; i_v < end_support; ++i_v) {
            int64_t v_indx = v_ci + i_v;
            int64_t v_oi = (int64_t)oversampling * i_v + v_coi;
            if (v_oi < 0) v_oi = -v_oi;
            double conv_v = cgk_1D[v_oi];
            int64_t v_indx_conj = v_ci_conj + i_v;
            for (int i_u = start_support; i_u < end_support; ++i_u) {
              int64_t u_indx = u_ci + i_u;
              int64_t u_oi = (int64_t)oversampling * i_u + u_coi;
              if (u_oi < 0) u_oi = -u_oi;
              double conv_u = cgk_1D[u_oi];
              double conv = conv_u * conv_v;
              int64_t cell = u_indx * n_v + v_indx;
              if (complex_grid) {
                g[2 * cell] = g[2 * cell] + conv * wd_re;
                g[2 * cell + 1] = g[2 * cell + 1] + conv * wd_im;
              } else {
                g[cell] = g[cell] + conv * wd_re;
              }
              norm = norm + conv;
              if (do_imaging_weight) {
                int64_t u_indx_conj = u_ci_conj + i_u;
                if (u_indx_conj >= 0 && u_indx_conj < n_u && v_indx_conj >= 0 &&
                    v_indx_conj < n_v) {
                  int64_t cc = u_indx_conj * n_v + v_indx_conj;
                  if (complex_grid) {
                    g[2 * cc] = g[2 * cc] + conv * wd_re;
                    g[2 * cc + 1] = g[2 * cc + 1] + conv * wd_im;
                  } else {
                    g[cc] = g[cc] + conv * wd_re;
                  }
                }
              }
            }
          }
          double *sw = sum_weight + a_chan * n_img_pol + a_pol;
          *sw = *sw + sel_weight * norm;
          if (do_imaging_weight) *sw = *sw + sel_weight * norm;
        }
      }
    }
  }
  free(uv_scale_u);
  free(uv_scale_v);
}

/* np.round(x): round half to even (matches the default rounding mode). */
static double round_half_even(double x) { return nearbyint(x); }

/* dgrid_numba, core/imaging/imaging_utils/standard_degrid.py:388-485, with
 * dphase = 0 so the phasor is exactly 1 (:538, :452-453).
 *
 * values complex128 [nt][nb][nvischan][nvispol], accumulated (+=).
 * flag   uint8 [nt][nb][nvischan][nvispol].
 * grid   complex128 [nchan_g][npol_g][g2][g3]; the reference reads
 *        nchan, npol, ny, nx = grid.shape (:407), so the u bound uses g3 and
 *        the v bound uses g2 (:457-462); both are also checked against the
 *        axis they index, which only differs for non-square grids where the
 *        reference itself would read out of bounds.
 * Loop order time -> baseline -> pol -> chan (:416-423), taps ix outer, iy
 * inner (:465-481); pos = scale*uvw*freq/c + offset evaluated left to right
 * (:435-440); loc = round_half_even(pos) (:442-443);
 * off = round_half_even((loc - pos) * sampling) (:445-450); the sum is divided
 * by the kernel-weight sum norm (:482-485). */
void ora_dgrid(const double *uvw, double *values, const uint8_t *flag,
               int nt, int nb, int nvischan, int nvispol, const double *scale,
               const double *offset, const double *grid, int nchan_g,
               int npol_g, int g2, int g3, const double *freq, double c,
               int support, int sampling, const double *convFunc,
               const int64_t *chanmap, const int64_t *polmap) {
  const int supp_centre = support / 2;
  const int supp_beg = -supp_centre;
  const int supp_end = supp_beg + support;
  const int64_t ny = g2, nx = g3;
  for (int t = 0; t < nt; ++t) {
    for (int b = 0; b < nb; ++b) {
      const int64_t row = (int64_t)t * nb + b;
      for (int ipol = 0; ipol < nvispol; ++ipol) {
        int64_t apol = polmap[ipol];
        if (!(0 <= apol && apol < npol_g)) continue;
        for (int ichan = 0; ichan < nvischan; ++ichan) {
          int64_t achan = chanmap[ichan];
          const int64_t vi = ((row * nvischan) + ichan) * nvispol + ipol;
          if (!(0 <= achan && achan < nchan_g && !flag[vi])) continue;
          double pos0 = scale[0] * uvw[row * 3 + 0] * freq[ichan] / c + offset[0];
          double pos1 = scale[1] * uvw[row * 3 + 1] * freq[ichan] / c + offset[1];
          double r0 = round_half_even(pos0), r1 = round_half_even(pos1);
          if (!(r0 > -2.0e9 && r0 < 2.0e9 && r1 > -2.0e9 && r1 < 2.0e9)) continue;
          int32_t loc0 = (int32_t)r0, loc1 = (int32_t)r1;
          int32_t off0 = (int32_t)round_half_even(((double)loc0 - pos0) * (double)sampling);
          int32_t off1 = (int32_t)round_half_even(((double)loc1 - pos1) * (double)sampling);
          int64_t x0 = loc0, y0 = loc1;
          if (!((x0 + supp_beg >= 0) && ((x0 + supp_end - 1) < nx) &&
                (y0 + supp_beg >= 0) && ((y0 + supp_end - 1) < ny)))
            continue;
          if (!((x0 + supp_end - 1) < g2 && (y0 + supp_end - 1) < g3)) continue;
          const double *g = grid + ((achan * npol_g) + apol) * (int64_t)g2 * g3 * 2;
          double nv_re = 0.0, nv_im = 0.0, norm = 0.0;
          for (int ix = supp_beg; ix < supp_end; ++ix) {
            int64_t il1 = (int64_t)sampling * ix + off0;
            if (il1 < 0) il1 = -il1;
            double wtx = convFunc[il1];
            for (int iy = supp_beg; iy < supp_end; ++iy) {
              int64_t il2 = (int64_t)sampling * iy + off1;
              if (il2 < 0) il2 = -il2;
              double wty = convFunc[il2];
              double wt = wtx * wty;
              norm += wt;
              int64_t cell = (x0 + ix) * g3 + (y0 + iy);
              nv_re += wt * g[2 * cell];
              nv_im += wt * g[2 * cell + 1];
            }
          }
          if (norm != 0.0) {
            values[2 * vi] += nv_re / norm;
            values[2 * vi + 1] += nv_im / norm;
          }
        }
      }
    }
  }
}

/* Post-FFT normalisation of grid2image_spheroid_ms4,
 * core/imaging/imaging_utils/standard_grid.py:696-708:
 * out = real_part / corrTerm * nx * ny / sumwt, evaluated left to right. */
void ora_normalise(const double *real_part, const double *corr, int64_t nplane,
                   int nx, int ny, const double *sumwt, double *out) {
  const int64_t np = (int64_t)nx * ny;
  for (int64_t p = 0; p < nplane; ++p)
    for (int64_t k = 0; k < np; ++k)
      out[p * np + k] =
          real_part[p * np + k] / corr[k] * (double)nx * (double)ny / sumwt[p];
}

/* Per-sample geometry of standard_grid_jit (standard_grid.py:296-352):
 * out[0..3] = u centre, v centre, u offset index, v offset index; returns 1
 * if the sample passes the bounds test, 0 otherwise (out untouched). */
int ora_grid_geometry(double uvw_u, double uvw_v, double freq, double dl_u, double dl_v,
                      int n_u, int n_v, int support, int oversampling, int64_t out[4]) {
  const double c = 299792458.0;
  const int sc = support / 2;
  double su = -(freq * dl_u * (double)n_u) / c;
  double sv = -(freq * dl_v * (double)n_v) / c;
  double u = uvw_u * su, v = uvw_v * sv;
  if (isnan(u) || isnan(v)) return 0;
  double u_pos = u + (double)(n_u / 2), v_pos = v + (double)(n_v / 2);
  double xu = u_pos + 0.5, xv = v_pos + 0.5;
  if (!(xu > -9.0e18 && xu < 9.0e18 && xv > -9.0e18 && xv < 9.0e18)) return 0;
  int64_t cu = (int64_t)xu, cv = (int64_t)xv;
  if (!((cu + sc < n_u) && (cv + sc < n_v) && (cu - sc >= 0) && (cv - sc >= 0))) return 0;
  out[0] = cu;
  out[1] = cv;
  out[2] = (int64_t)floor(((double)cu - u_pos) * (double)oversampling + 0.5);
  out[3] = (int64_t)floor(((double)cv - v_pos) * (double)oversampling + 0.5);
  return 1;
}

/* Per-sample geometry of dgrid_numba / sgrid_numba
 * (standard_degrid.py:367-385, 435-450): pos, loc (round half even) and
 * off.  pos_out[2], loc_off[4] = loc0, loc1, off0, off1. */
void ora_degrid_geometry(const double uvw[3], double freq, const double scale[2],
                         const double offset[2], int sampling, double pos_out[2],
                         int64_t loc_off[4]) {
  const double c = 299792458.0;
  double p0 = scale[0] * uvw[0] * freq / c + offset[0];
  double p1 = scale[1] * uvw[1] * freq / c + offset[1];
  pos_out[0] = p0;
  pos_out[1] = p1;
  int32_t l0 = (int32_t)round_half_even(p0), l1 = (int32_t)round_half_even(p1);
  loc_off[0] = l0;
  loc_off[1] = l1;
  loc_off[2] = (int32_t)round_half_even(((double)l0 - p0) * (double)sampling);
  loc_off[3] = (int32_t)round_half_even(((double)l1 - p1) * (double)sampling);
}

/* Transpose of ora_dgrid (no reference function: the measurement operator's
 * backward).  For every unflagged sample with dgrid_numba's geometry,
 * grid[x0+ix, y0+iy] += wt * values / norm with wt and norm exactly as in
 * ora_dgrid, so <dgrid(g), y> = <g, ora_grid_adjoint(y)> to rounding. */
void ora_grid_adjoint(const double *uvw, const double *values, const uint8_t *flag,
                      int nt, int nb, int nvischan, int nvispol, const double *scale,
                      const double *offset, double *grid, int nchan_g, int npol_g, int g2,
                      int g3, const double *freq, double c, int support, int sampling,
                      const double *convFunc, const int64_t *chanmap, const int64_t *polmap) {
  const int supp_centre = support / 2;
  const int supp_beg = -supp_centre;
  const int supp_end = supp_beg + support;
  const int64_t ny = g2, nx = g3;
  for (int t = 0; t < nt; ++t) {
    for (int b = 0; b < nb; ++b) {
      const int64_t row = (int64_t)t * nb + b;
      for (int ipol = 0; ipol < nvispol; ++ipol) {
        int64_t apol = polmap[ipol];
        if (!(0 <= apol && apol < npol_g)) continue;
        for (int ichan = 0; ichan < nvischan; ++ichan) {
          int64_t achan = chanmap[ichan];
          const int64_t vi = ((row * nvischan) + ichan) * nvispol + ipol;
          if (!(0 <= achan && achan < nchan_g && !(flag && flag[vi]))) continue;
          double pos0 = scale[0] * uvw[row * 3 + 0] * freq[ichan] / c + offset[0];
          double pos1 = scale[1] * uvw[row * 3 + 1] * freq[ichan] / c + offset[1];
          double r0 = round_half_even(pos0), r1 = round_half_even(pos1);
          if (!(r0 > -2.0e9 && r0 < 2.0e9 && r1 > -2.0e9 && r1 < 2.0e9)) continue;
          int32_t loc0 = (int32_t)r0, loc1 = (int32_t)r1;
          int32_t off0 = (int32_t)round_half_even(((double)loc0 - pos0) * (double)sampling);
          int32_t off1 = (int32_t)round_half_even(((double)loc1 - pos1) * (double)sampling);
          int64_t x0 = loc0, y0 = loc1;
          if (!((x0 + supp_beg >= 0) && ((x0 + supp_end - 1) < nx) && (y0 + supp_beg >= 0) &&
                ((y0 + supp_end - 1) < ny)))
            continue;
          if (!((x0 + supp_end - 1) < g2 && (y0 + supp_end - 1) < g3)) continue;
          double *g = grid + ((achan * npol_g) + apol) * (int64_t)g2 * g3 * 2;
          double norm = 0.0;
          for (int ix = supp_beg; ix < supp_end; ++ix) {
            int64_t il1 = (int64_t)sampling * ix + off0;
            if (il1 < 0) il1 = -il1;
            for (int iy = supp_beg; iy < supp_end; ++iy) {
              int64_t il2 = (int64_t)sampling * iy + off1;
              if (il2 < 0) il2 = -il2;
              norm += convFunc[il1] * convFunc[il2];
            }
          }
          if (norm == 0.0) continue;
          double yr = values[2 * vi] / norm, yi = values[2 * vi + 1] / norm;
          for (int ix = supp_beg; ix < supp_end; ++ix) {
            int64_t il1 = (int64_t)sampling * ix + off0;
            if (il1 < 0) il1 = -il1;
            double wtx = convFunc[il1];
            for (int iy = supp_beg; iy < supp_end; ++iy) {
              int64_t il2 = (int64_t)sampling * iy + off1;
              if (il2 < 0) il2 = -il2;
              double wt = wtx * convFunc[il2];
              int64_t cell = (x0 + ix) * g3 + (y0 + iy);
              g[2 * cell] += wt * yr;
              g[2 * cell + 1] += wt * yi;
            }
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------
 * Hogbom CLEAN, deconvolvers/hogbom/src/hclean.cpp (the reference's own C++,
 * a restatement of the AIPS++ Fortran).  Images are [ny][nx] row-major.
 *
 * ora_maximg_T: hclean.cpp:51-81 -- min and max over the (masked) image,
 * starting from +-1e20.
 * ora_hogbom_T: hclean.cpp:106-210 without the progress/stop callbacks:
 * every iteration finds the first pixel of largest |residual| inside the
 * box (row-major scan, strict >, so |r| = 0 or NaN never wins and the
 * previous peak position is kept), stops when that is below the threshold
 * (scaled by 2^((iter - siter)/cspeedup) when cspeedup > 0), adds
 * gain * peak to the model and subtracts gain * peak * psf centred at
 * (nx/2, ny/2) shifted to the peak.  Returns the last iteration number.
 * ------------------------------------------------------------------------ */
#define ORA_HOGBOM(T, SFX, ABS, POW)                                                    \
  void ora_maximg_##SFX(const T *img, int domask, const T *mask, int nx, int ny,        \
                        T *fmin, T *fmax) {                                             \
    T lo = (T)1e20, hi = (T)-1e20;                                                      \
    for (int iy = 0; iy < ny; ++iy)                                                     \
      for (int ix = 0; ix < nx; ++ix)                                                   \
        if (domask == 0 || mask[iy * nx + ix] > (T)0.5) {                               \
          T w = img[iy * nx + ix];                                                      \
          if (w > hi) hi = w;                                                           \
          if (w < lo) lo = w;                                                           \
        }                                                                               \
    *fmin = lo;                                                                         \
    *fmax = hi;                                                                         \
  }                                                                                     \
  int ora_hogbom_##SFX(T *model, T *resid, const T *psf, int domask, const T *mask,     \
                       int nx, int ny, int xbeg, int xend, int ybeg, int yend,          \
                       int niter, int siter, T gain, T thres, T cspeedup, int *peak) {  \
    int px = 0, py = 0, iter;                                                           \
    T maxval = (T)0;                                                                    \
    for (iter = siter; iter < niter; ++iter) {                                          \
      T absval = (T)0;                                                                  \
      for (int iy = ybeg; iy < yend; ++iy)                                              \
        for (int ix = xbeg; ix < xend; ++ix)                                            \
          if (domask == 0 || mask[iy * nx + ix] > (T)0.5) {                             \
            T v = ABS(resid[iy * nx + ix]);                                             \
            if (v > absval) {                                                           \
              px = ix;                                                                  \
              py = iy;                                                                  \
              absval = v;                                                               \
            }                                                                           \
          }                                                                             \
      maxval = resid[py * nx + px];                                                     \
      T cthres = cspeedup > (T)0 ? thres * POW((T)2, (T)(iter - siter) / cspeedup)     \
                                 : thres;                                               \
      if (absval < cthres) break;                                                       \
      int x1 = px - nx / 2 > 0 ? px - nx / 2 : 0;                                       \
      int y1 = py - ny / 2 > 0 ? py - ny / 2 : 0;                                       \
      int x2 = px + nx / 2 - 1 < nx - 1 ? px + nx / 2 - 1 : nx - 1;                     \
      int y2 = py + ny / 2 - 1 < ny - 1 ? py + ny / 2 - 1 : ny - 1;                     \
      T pv = gain * maxval;                                                             \
      model[py * nx + px] += pv;                                                        \
      for (int iy = y1; iy <= y2; ++iy)                                                 \
        for (int ix = x1; ix <= x2; ++ix) {                                             \
          int qx = nx / 2 + ix - px, qy = ny / 2 + iy - py;                             \
          if (qx < 0 || qx >= nx || qy < 0 || qy >= ny) continue;                       \
          resid[iy * nx + ix] -= pv * psf[qy * nx + qx];                                \
        }                                                                               \
    }                                                                                   \
    if (iter > niter) iter = niter;                                                     \
    if (peak) {                                                                         \
      peak[0] = px;                                                                     \
      peak[1] = py;                                                                     \
    }                                                                                   \
    return iter;                                                                        \
  }

ORA_HOGBOM(double, f64, fabs, pow)
ORA_HOGBOM(float, f32, fabsf, powf)

/* ------------------------------------------------------------------------
 * Imaging weights, core/imaging/imaging_weighting/grid_imaging_weights.py.
 *
 * ora_grid_imaging_weights: grid_imaging_weights_jit (:105-206).  For each
 * (time, baseline, channel) in that order: u = uvw_u * -(f*dl*N)/c (likewise
 * v); skip NaN; nearest pixel int(u + N//2 + 0.5) (truncation toward zero)
 * and the conjugate int(-u + N//2 + 0.5); if the direct pixel is inside the
 * grid and the weight (polarisation 0) is not NaN, add the weight at both
 * pixels and 2 * weight to sum_weight.  The reference does not bound the
 * conjugate pixel (for a sample on row or column 0 of an even grid it is N,
 * one past the end); this restatement, like the GPU, skips such a
 * conjugate update and keeps the rest.
 * grid [n_chan][n_u][n_v], sum_weight [n_chan], weight [nt][nb][nchan][npol].
 * ------------------------------------------------------------------------ */
void ora_grid_imaging_weights(double *grid, double *sum_weight, const double *uvw, int nt,
                              int nb, const double *freq, int nchan, const double *weight,
                              int npol_w, int n_u, int n_v, double cell_u, double cell_v) {
  const double c = 299792458.0;
  const int cu = n_u / 2, cv = n_v / 2;
  for (int t = 0; t < nt; ++t)
    for (int b = 0; b < nb; ++b)
      for (int ch = 0; ch < nchan; ++ch) {
        const double su = -(freq[ch] * cell_u * (double)n_u) / c;
        const double sv = -(freq[ch] * cell_v * (double)n_v) / c;
        const int64_t row = (int64_t)t * nb + b;
        const double u = uvw[row * 3] * su, v = uvw[row * 3 + 1] * sv;
        if (isnan(u) || isnan(v)) continue;
        const double up = u + (double)cu, vp = v + (double)cv;
        const double upc = -u + (double)cu, vpc = -v + (double)cv;
        if (!(fabs(up) < 1e9 && fabs(vp) < 1e9)) continue;
        const int iu = (int)(up + 0.5), iv = (int)(vp + 0.5);
        if (!(iu < n_u && iv < n_v && iu >= 0 && iv >= 0)) continue;
        const double w = weight[(row * nchan + ch) * npol_w];
        if (isnan(w)) continue;
        double *g = grid + (int64_t)ch * n_u * n_v;
        g[(int64_t)iu * n_v + iv] = g[(int64_t)iu * n_v + iv] + w;
        const int iuc = (int)(upc + 0.5), ivc = (int)(vpc + 0.5);
        if (iuc >= 0 && iuc < n_u && ivc >= 0 && ivc < n_v)
          g[(int64_t)iuc * n_v + ivc] = g[(int64_t)iuc * n_v + ivc] + w;
        sum_weight[ch] = sum_weight[ch] + 2.0 * w;
      }
}

/* ora_degrid_imaging_weights: degrid_imaging_weights_jit (:274-443) with
 * the single-polarisation map of degrid_imaging_weights (:260): imaging
 * weight = data weight where the sample's pixel is inside the grid (0
 * elsewhere, the output starts at zero), divided by briggs[0][ch] * g +
 * briggs[1][ch] when the data weight is finite and non-zero and the grid
 * value g at the pixel is finite and non-zero.  Returns 1 when a Briggs
 * factor used is NaN (the reference raises ValueError), else 0. */
int ora_degrid_imaging_weights(double *imaging_weight, const double *grid, const double *briggs,
                               const double *uvw, int nt, int nb, const double *freq, int nchan,
                               const double *weight, int n_u, int n_v, double cell_u,
                               double cell_v) {
  const double c = 299792458.0;
  const int cu = n_u / 2, cv = n_v / 2;
  for (int t = 0; t < nt; ++t)
    for (int b = 0; b < nb; ++b)
      for (int ch = 0; ch < nchan; ++ch) {
        const double su = -(freq[ch] * cell_u * (double)n_u) / c;
        const double sv = -(freq[ch] * cell_v * (double)n_v) / c;
        const int64_t row = (int64_t)t * nb + b;
        const double u = uvw[row * 3] * su, v = uvw[row * 3 + 1] * sv;
        if (isnan(u) || isnan(v)) continue;
        const double up = u + (double)cu, vp = v + (double)cv;
        if (!(fabs(up) < 1e9 && fabs(vp) < 1e9)) continue;
        const int iu = (int)(up + 0.5), iv = (int)(vp + 0.5);
        if (!(iu < n_u && iv < n_v && iu >= 0 && iv >= 0)) continue;
        const int64_t k = row * nchan + ch;
        const double w = weight[k];
        imaging_weight[k] = w;
        if (isnan(w) || w == 0.0) continue;
        const double g = grid[((int64_t)ch * n_u + iu) * n_v + iv];
        if (isnan(g) || g == 0.0) continue;
        if (isnan(briggs[ch]) || isnan(briggs[nchan + ch])) return 1;
        const double denom = briggs[ch] * g + briggs[nchan + ch];
        imaging_weight[k] = imaging_weight[k] / denom;
      }
  return 0;
}

/* mosaic_grid_jit, core/imaging/imaging_utils/mosaic_grid.py:306-486: the
 * A-projection / mosaic gridder with a non-separable oversampled complex
 * convolution function per (cf baseline, cf channel, cf polarisation),
 * already multiplied by phase_gradient[0] (mosaic_grid.py:359, done by the
 * caller, as conv_kernel * phase_gradient[0, :, :]).  Complex grids only
 * (numba cannot store the complex product into a real grid).  cfpg is
 * [ncb][ncc][ncp][cu_n][cv_n] complex, weight_support [ncb][ncc][ncp][2].
 * The reference does not bound the cf indices; a tap outside the cf array is
 * skipped here (no in-range input reaches it). */
void ora_mosaic_grid(double *grid, double *sum_weight, int do_psf, const double *vis,
                     int nvpol, const double *uvw, int nt, int nb, const double *freq_chan,
                     int nchan, const int64_t *chan_map, const int64_t *pol_map, int npol,
                     const int64_t *cf_baseline_map, const int64_t *cf_chan_map,
                     const int64_t *cf_pol_map, const double *imaging_weight,
                     const double *cfpg, int ncc, int ncp, int cu_n, int cv_n, int n_img_pol,
                     int n_u, int n_v, double dl_u, double dl_v, const int64_t *weight_support,
                     int64_t max_support, int ov_u, int ov_v) {
  const double c = 299792458.0;
  const int64_t uc = n_u / 2, vc = n_v / 2;
  const int64_t conv_v_center = cv_n / 2, conv_u_center = cu_n / 2;
  const int64_t plane = (int64_t)n_u * n_v;
  for (int t = 0; t < nt; ++t)
    for (int b = 0; b < nb; ++b) {
      const int64_t cf_baseline = cf_baseline_map[b];
      for (int ch = 0; ch < nchan; ++ch) {
        const int64_t cf_chan = cf_chan_map[ch];
        const int64_t a_chan = chan_map[ch];
        const double su = -(freq_chan[ch] * dl_u * (double)n_u) / c;  /* :336-338 */
        const double sv = -(freq_chan[ch] * dl_v * (double)n_v) / c;
        const int64_t row = (int64_t)t * nb + b;
        const double u = uvw[row * 3 + 0] * su, v = uvw[row * 3 + 1] * sv;
        if (isnan(u) || isnan(v)) continue;
        const double u_pos = u + (double)uc, v_pos = v + (double)vc;
        const double xu = u_pos + 0.5, xv = v_pos + 0.5;
        if (!(xu > -9.0e18 && xu < 9.0e18 && xv > -9.0e18 && xv < 9.0e18)) continue;
        const int64_t u_ci = (int64_t)xu, v_ci = (int64_t)xv;
        if (!((u_ci + max_support < n_u) && (v_ci + max_support < n_v) &&
              (u_ci - max_support >= 0) && (v_ci - max_support >= 0)))
          continue;
        const double u_offset = (double)u_ci - u_pos;
        const int64_t u_coi = (int64_t)floor(u_offset * (double)ov_u + 0.5) + conv_u_center;
        const double v_offset = (double)v_ci - v_pos;
        const int64_t v_coi = (int64_t)floor(v_offset * (double)ov_v + 0.5) + conv_v_center;
        for (int ip = 0; ip < npol; ++ip) {
          const int64_t k = (row * nchan + ch) * npol + ip;
          const double iw = imaging_weight[k];
          double wd_re, wd_im;
          if (do_psf) {
            wd_re = iw;
            wd_im = 0.0;
          } else {
            const int64_t vi = ((row * nchan + ch) * nvpol + ip) * 2;
            wd_re = vis[vi] * iw;
            wd_im = vis[vi + 1] * iw;
          }
          if (isnan(wd_re) || isnan(wd_im)) continue;
          if (wd_re == 0.0 && wd_im == 0.0) continue;
          const int64_t cf_pol = cf_pol_map[ip];
          const int64_t a_pol = pol_map[ip];
          const int64_t *sup = weight_support + ((cf_baseline * ncc + cf_chan) * ncp + cf_pol) * 2;
          const int64_t support_u = sup[0], support_v = sup[1];
          const int64_t scu = support_u / 2, scv = support_v / 2;
          const double *cf = cfpg + (((cf_baseline * ncc + cf_chan) * ncp + cf_pol) *
                                     (int64_t)cu_n * cv_n) * 2;
          double *g = grid + ((a_chan * n_img_pol) + a_pol) * plane * 2;
          double n_re = 0.0, n_im = 0.0;
          for (int64_t i_v = -scv; i_v < support_v - scv; ++i_v) {
            const int64_t v_indx = v_ci + i_v;
            const int64_t cf_v = (int64_t)ov_v * i_v + v_coi;
            for (int64_t i_u = -scu; i_u < support_u - scu; ++i_u) {
              const int64_t u_indx = u_ci + i_u;
              const int64_t cf_u = (int64_t)ov_u * i_u + u_coi;
              if (cf_u < 0 || cf_u >= cu_n || cf_v < 0 || cf_v >= cv_n) continue;
              const double c_re = cf[(cf_u * cv_n + cf_v) * 2];
              const double c_im = cf[(cf_u * cv_n + cf_v) * 2 + 1];
              const int64_t cell = u_indx * n_v + v_indx;
              g[2 * cell] = g[2 * cell] + (c_re * wd_re - c_im * wd_im);
              g[2 * cell + 1] = g[2 * cell + 1] + (c_re * wd_im + c_im * wd_re);
              n_re = n_re + c_re;
              n_im = n_im + c_im;
            }
          }
          double *sw = sum_weight + a_chan * n_img_pol + a_pol;
          if (do_psf)
            *sw = *sw + iw * n_re;
          else
            *sw = *sw + iw * (n_re * n_re - n_im * n_im);
        }
      }
    }
}
