// spray_oracle_eb.h -- CPU restatement of the embedded-boundary parts of the
// droplet update.
//
// TEST INFRASTRUCTURE (see spray_oracle.h).  Restates:
//   Source/PP_Spray/SprayInterpolation.H:201-634  f / dfdxi / dfdeta / dfdzeta,
//     get_interp_mapping (Newton inversion of the trilinear map through the
//     cut-cell centroids), eb_interp (inverse-distance weights near covered
//     cells, the EB or small cells; a droplet behind the EB moved to the
//     boundary centroid of the closest cell)
//   Source/PP_Spray/WallFunctions.H:124-157       check_wall, EB part
//   Source/PP_Spray/SprayParticles.cpp:414-418    deposit scaled by 1/vfrac in
//                                                  single-valued cells
// EB data in AMReX's layout: EBCellFlag words (type in bits 0-1: 0 regular,
// 1 single-valued, 2 multi-valued, 3 covered; neighbour (i,j,k) connected
// when bit 5 + (i+1) + 3(j+1) + 9(k+1) is set), volume fraction, cell and
// boundary centroids relative to the cell centre in cell units, boundary
// normal; FABs over one box, component-major.
//
// Deviations (documented): a stencil without an uncovered cell of large
// enough volume fraction (the reference then indexes indx_array[-1]) is an
// error, PELEMP_E_EB_STENCIL.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

#include "../include/pelemp_spray.h"

namespace oracle {

struct EB {
  const uint32_t* flags = nullptr;
  const double* vfrac = nullptr;
  const double* ccent = nullptr;
  const double* bcent = nullptr;
  const double* bnorm = nullptr;
  int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  bool on = false;

  int64_t off(int i, int j, int k) const {
    const int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    return (static_cast<int64_t>(k - lo[2]) * ny + (j - lo[1])) * nx + (i - lo[0]);
  }
  int64_t cstride() const {
    return static_cast<int64_t>(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  }
  bool has(int i, int j, int k) const {
    return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
  }
  // cells outside the box count as regular
  uint32_t flag(int i, int j, int k) const { return has(i, j, k) ? flags[off(i, j, k)] : 0x7FFFFFE0u; }
  bool regular(int i, int j, int k) const { return (flag(i, j, k) & 3u) == 0u; }
  bool single(int i, int j, int k) const { return (flag(i, j, k) & 3u) == 1u; }
  bool covered(int i, int j, int k) const { return (flag(i, j, k) & 3u) == 3u; }
  bool connected(int i, int j, int k, int di, int dj, int dk) const {
    return (flag(i, j, k) >> (5 + (di + 1) + 3 * (dj + 1) + 9 * (dk + 1))) & 1u;
  }
  // outside the box: a full regular cell
  double vf(int i, int j, int k) const { return has(i, j, k) ? vfrac[off(i, j, k)] : 1.; }
  double cc(int i, int j, int k, int d) const { return has(i, j, k) ? ccent[off(i, j, k) + d * cstride()] : 0.; }
  double bc(int i, int j, int k, int d) const { return has(i, j, k) ? bcent[off(i, j, k) + d * cstride()] : 0.; }
  double bn(int i, int j, int k, int d) const { return has(i, j, k) ? bnorm[off(i, j, k) + d * cstride()] : 0.; }
};

// the trilinear map through the stencil nodes minus pos, component dir
// (SprayInterpolation.H:207-240), and its partial derivatives
inline double eb_f(int dir, const double pos[3], const double nodes[8][3], double xi, double eta, double zeta) {
  const double a0 = (nodes[0][dir] - pos[dir]);
  const double a1 = (nodes[1][dir] - nodes[0][dir]);
  const double a2 = (nodes[2][dir] - nodes[0][dir]);
  const double a4 = (nodes[0][dir] - nodes[1][dir] + nodes[3][dir] - nodes[2][dir]);
  const double a3 = (nodes[4][dir] - nodes[0][dir]);
  const double a5 = (nodes[0][dir] - nodes[1][dir] - nodes[4][dir] + nodes[5][dir]);
  const double a6 = (nodes[0][dir] - nodes[2][dir] - nodes[4][dir] + nodes[6][dir]);
  const double a7 = (nodes[1][dir] - nodes[3][dir] + nodes[2][dir] + nodes[4][dir] - nodes[5][dir] + nodes[7][dir] -
                     nodes[6][dir] - nodes[0][dir]);
  return a0 + a1 * xi + a2 * eta + a3 * zeta + a4 * xi * eta + a5 * xi * zeta + a6 * eta * zeta +
         a7 * xi * eta * zeta;
}
inline double eb_dfdxi(int dir, const double nodes[8][3], double eta, double zeta) {
  const double a1 = (nodes[1][dir] - nodes[0][dir]);
  const double a4 = (nodes[0][dir] - nodes[1][dir] + nodes[3][dir] - nodes[2][dir]);
  const double a5 = (nodes[0][dir] - nodes[1][dir] - nodes[4][dir] + nodes[5][dir]);
  const double a7 = (nodes[1][dir] - nodes[3][dir] + nodes[2][dir] + nodes[4][dir] - nodes[5][dir] + nodes[7][dir] -
                     nodes[6][dir] - nodes[0][dir]);
  return a1 + a4 * eta + a5 * zeta + a7 * eta * zeta;
}
inline double eb_dfdeta(int dir, const double nodes[8][3], double xi, double zeta) {
  const double a2 = (nodes[2][dir] - nodes[0][dir]);
  const double a4 = (nodes[0][dir] - nodes[1][dir] + nodes[3][dir] - nodes[2][dir]);
  const double a6 = (nodes[0][dir] - nodes[2][dir] - nodes[4][dir] + nodes[6][dir]);
  const double a7 = (nodes[1][dir] - nodes[3][dir] + nodes[2][dir] + nodes[4][dir] - nodes[5][dir] + nodes[7][dir] -
                     nodes[6][dir] - nodes[0][dir]);
  return a2 + a4 * xi + a6 * zeta + a7 * xi * zeta;
}
inline double eb_dfdzeta(int dir, const double nodes[8][3], double xi, double eta) {
  const double a3 = (nodes[4][dir] - nodes[0][dir]);
  const double a5 = (nodes[0][dir] - nodes[1][dir] - nodes[4][dir] + nodes[5][dir]);
  const double a6 = (nodes[0][dir] - nodes[2][dir] - nodes[4][dir] + nodes[6][dir]);
  const double a7 = (nodes[1][dir] - nodes[3][dir] + nodes[2][dir] + nodes[4][dir] - nodes[5][dir] + nodes[7][dir] -
                     nodes[6][dir] - nodes[0][dir]);
  return a3 + a5 * xi + a6 * eta + a7 * xi * eta;
}

// get_interp_mapping (SprayInterpolation.H:320-393): Newton with Cramer's
// rule, at most 10 steps, until the largest change is at most 1e-3
inline void get_interp_mapping(const double pos[3], const double nodes[8][3], double& xi, double& eta,
                               double& zeta) {
  int lc = 0;
  double err = 1.0;
  while (err > 1.0e-3 && lc < 10) {
    const double f0 = eb_f(0, pos, nodes, xi, eta, zeta);
    const double f1 = eb_f(1, pos, nodes, xi, eta, zeta);
    const double df0dxi = eb_dfdxi(0, nodes, eta, zeta);
    const double df0deta = eb_dfdeta(0, nodes, xi, zeta);
    const double df1dxi = eb_dfdxi(1, nodes, eta, zeta);
    const double df1deta = eb_dfdeta(1, nodes, xi, zeta);
    const double f2 = eb_f(2, pos, nodes, xi, eta, zeta);
    const double df0dzeta = eb_dfdzeta(0, nodes, xi, eta);
    const double df1dzeta = eb_dfdzeta(1, nodes, xi, eta);
    const double df2dxi = eb_dfdxi(2, nodes, eta, zeta);
    const double df2deta = eb_dfdeta(2, nodes, xi, zeta);
    const double df2dzeta = eb_dfdzeta(2, nodes, xi, eta);
    const double detJ = df0dxi * (df1deta * df2dzeta - df1dzeta * df2deta) -
                        df0deta * (df1dxi * df2dzeta - df1dzeta * df2dxi) +
                        df0dzeta * (df1dxi * df2deta - df1deta * df2dxi);
    const double detJ_xi = f0 * (df1deta * df2dzeta - df1dzeta * df2deta) - df0deta * (f1 * df2dzeta - df1dzeta * f2) +
                           df0dzeta * (f1 * df2deta - df1deta * f2);
    const double detJ_eta = df0dxi * (f1 * df2dzeta - df1dzeta * f2) - f0 * (df1dxi * df2dzeta - df1dzeta * df2dxi) +
                            df0dzeta * (df1dxi * f2 - f1 * df2dxi);
    const double detJ_zeta = df0dxi * (df1deta * f2 - f1 * df2deta) - df0deta * (df1dxi * f2 - f1 * df2dxi) +
                             f0 * (df1dxi * df2deta - df1deta * df2dxi);
    const double new_xi = xi - detJ_xi / detJ;
    const double new_eta = eta - detJ_eta / detJ;
    const double new_zeta = zeta - detJ_zeta / detJ;
    lc += 1;
    err = std::max(std::fabs(xi - new_xi), std::max(std::fabs(eta - new_eta), std::fabs(zeta - new_zeta)));
    xi = new_xi;
    eta = new_eta;
    zeta = new_zeta;
  }
}

// eb_interp (SprayInterpolation.H:395-632).  Index space relative to the
// domain's lower corner (eb indices are these plus domain_lo).  Returns
// do_fe_interp; ijkc and pos may change; status set on an empty stencil.
// idx / w: the 8 cells and weights in the reference's order.
template <class TrilinearFn>
inline bool eb_interp(double pos[3], int ijkc[3], const int ijk[3], const pelemp_geom& g, const double dxi[3],
                      const EB& eb, double min_eb_vfrac, int idx[8][3], double w[8], TrilinearFn trilinear,
                      int& status) {
  const int* dl = g.domain_lo;
  const int ip = ijkc[0], jp = ijkc[1], kp = ijkc[2];
  double normpos[3], normposp[3];
  for (int d = 0; d < 3; ++d) {
    normpos[d] = (pos[d] - g.plo[d]) * dxi[d];
    normposp[d] = normpos[d] - 0.5;
  }
  const int ic = ijk[0], jc = ijk[1], kc = ijk[2];
  const int ks = kc - 1, ke = kc;
  bool do_fe_interp = false;
  for (int kk = ks; kk <= ke; kk++)
    for (int jj = jc - 1; jj <= jc; jj++)
      for (int ii = ic - 1; ii <= ic; ii++)
        if (!eb.regular(ii + dl[0], jj + dl[1], kk + dl[2])) do_fe_interp = true;
  if (!do_fe_interp) {
    trilinear(idx, w);
    return false;
  }
  const double tolerance = std::numeric_limits<double>::epsilon();
  double par_dot_EB = 2.;
  double cent_dot_EB = 1.;
  const int Ip = ip + dl[0], Jp = jp + dl[1], Kp = kp + dl[2];
  if (eb.single(Ip, Jp, Kp)) {
    double normal[3];
    for (int d = 0; d < 3; ++d) normal[d] = -eb.bn(Ip, Jp, Kp, d);
    par_dot_EB = (normposp[0] - (ip + eb.bc(Ip, Jp, Kp, 0))) * normal[0] +
                 (normposp[1] - (jp + eb.bc(Ip, Jp, Kp, 1))) * normal[1] +
                 (normposp[2] - (kp + eb.bc(Ip, Jp, Kp, 2))) * normal[2];
    cent_dot_EB = (eb.cc(Ip, Jp, Kp, 0) - eb.bc(Ip, Jp, Kp, 0)) * normal[0] +
                  (eb.cc(Ip, Jp, Kp, 1) - eb.bc(Ip, Jp, Kp, 1)) * normal[1] +
                  (eb.cc(Ip, Jp, Kp, 2) - eb.bc(Ip, Jp, Kp, 2)) * normal[2];
  }
  int covered = 0;
  for (int kk = ks; kk <= ke; kk++)
    for (int jj = jc - 1; jj <= jc; jj++)
      for (int ii = ic - 1; ii <= ic; ii++)
        if (!eb.connected(Ip, Jp, Kp, ii - ip, jj - jp, kk - kp)) covered += 1;
  const bool in_covered = eb.covered(Ip, Jp, Kp);
  const bool behind_EB = (par_dot_EB < tolerance);
  const bool bad_part = in_covered || behind_EB;
  const bool near_covered = (covered > 0);
  const bool near_EB = (par_dot_EB < cent_dot_EB);
  const bool in_small_cell = (eb.vf(Ip, Jp, Kp) < min_eb_vfrac);
  const bool use_invw = near_covered || near_EB || in_small_cell || bad_part;
  if (use_invw) {
    int aindx = 0;
    int closest_aindx = -1;
    double sum_wt = 0.;
    double min_dis = 1.E12;
    bool at_cent = false;
    for (int kk = ks; kk <= ke; kk++) {
      for (int jj = jc - 1; jj <= jc; jj++) {
        const double y = static_cast<double>(jj);
        for (int ii = ic - 1; ii <= ic; ii++) {
          const double x = static_cast<double>(ii);
          idx[aindx][0] = ii;
          idx[aindx][1] = jj;
          idx[aindx][2] = kk;
          const int I = ii + dl[0], J = jj + dl[1], K = kk + dl[2];
          if (!eb.covered(I, J, K) && eb.vf(I, J, K) > min_eb_vfrac) {
            const double dlx = normposp[0] - (x + eb.cc(I, J, K, 0));
            const double dly = normposp[1] - (y + eb.cc(I, J, K, 1));
            const double dlz = normposp[2] - (static_cast<double>(kk) + eb.cc(I, J, K, 2));
            const double len = std::sqrt(dlx * dlx + dly * dly + dlz * dlz);
            if (len == 0.) {
              w[aindx] = 1.;
              at_cent = true;
            } else {
              w[aindx] = 1. / len;
              sum_wt += 1. / len;
            }
            if (len < min_dis) {
              closest_aindx = aindx;
              min_dis = len;
            }
          } else {
            w[aindx] = 0.;
          }
          aindx++;
        }
      }
    }
    if (closest_aindx < 0) {
      status = PELEMP_E_EB_STENCIL;
      return true;
    }
    if (at_cent) {
      for (aindx = 1; aindx < 8; ++aindx) w[aindx] = 0.;
    } else {
      for (aindx = 0; aindx < 8; ++aindx) w[aindx] /= sum_wt;
    }
    for (int d = 0; d < 3; ++d) ijkc[d] = idx[closest_aindx][d];
    if (bad_part) {
      const int I = ijkc[0] + dl[0], J = ijkc[1] + dl[1], K = ijkc[2] + dl[2];
      for (int dir = 0; dir < 3; ++dir)
        pos[dir] = (static_cast<double>(ijkc[dir]) + 0.5 + eb.bc(I, J, K, dir)) * g.dx[dir] + g.plo[dir];
    }
  } else {
    double nodes[8][3];
    int lc = 0;
    for (int kk = ks; kk <= ke; kk++) {
      for (int jj = jc - 1; jj <= jc; jj++) {
        for (int ii = ic - 1; ii <= ic; ii++) {
          const int I = ii + dl[0], J = jj + dl[1], K = kk + dl[2];
          nodes[lc][0] = (ii + 0.5 + eb.cc(I, J, K, 0));
          nodes[lc][1] = (jj + 0.5 + eb.cc(I, J, K, 1));
          nodes[lc][2] = (kk + 0.5 + eb.cc(I, J, K, 2));
          lc += 1;
        }
      }
    }
    double xi = normpos[0] - nodes[0][0];
    double eta = normpos[1] - nodes[0][1];
    double zeta = normpos[2] - nodes[0][2];
    get_interp_mapping(normpos, nodes, xi, eta, zeta);
    const int order[8][3] = {{-1, -1, -1}, {0, -1, -1}, {0, 0, -1}, {-1, 0, -1},
                             {-1, -1, 0},  {0, -1, 0},  {0, 0, 0},  {-1, 0, 0}};
    for (int a = 0; a < 8; ++a) {
      idx[a][0] = ic + order[a][0];
      idx[a][1] = jc + order[a][1];
      idx[a][2] = kc + order[a][2];
    }
    w[0] = 1. - xi - eta - zeta + xi * eta + xi * zeta + eta * zeta - xi * eta * zeta;
    w[1] = xi * (1. - eta - zeta + eta * zeta);
    w[2] = xi * eta - xi * eta * zeta;
    w[3] = eta * (1. - xi - zeta + xi * zeta);
    w[4] = zeta * (1. - xi - eta + xi * eta);
    w[5] = xi * zeta - xi * eta * zeta;
    w[6] = xi * eta * zeta;
    w[7] = eta * zeta - xi * eta * zeta;
  }
  return true;
}

// check_wall's EB part (WallFunctions.H:124-157): pos normalised, ijkc the
// new cell, ijkc_prev the cell of the gather (domain-relative).  Fills
// normal and the boundary point bcentv (cell units); returns wall_check.
inline bool eb_wall(const EB& eb, const pelemp_geom& g, const int ijkc[3], const int ijkc_prev[3], double normal[3],
                    double bcentv[3]) {
  const int* dl = g.domain_lo;
  const int I = ijkc[0] + dl[0], J = ijkc[1] + dl[1], K = ijkc[2] + dl[2];
  if (eb.single(I, J, K)) {
    for (int d = 0; d < 3; ++d) {
      normal[d] = -eb.bn(I, J, K, d);
      bcentv[d] = ijkc[d] + 0.5 + eb.bc(I, J, K, d);
    }
    return true;
  }
  if (eb.covered(I, J, K)) {
    const int Ip = ijkc_prev[0] + dl[0], Jp = ijkc_prev[1] + dl[1], Kp = ijkc_prev[2] + dl[2];
    if (eb.single(Ip, Jp, Kp)) {
      for (int d = 0; d < 3; ++d) {
        normal[d] = -eb.bn(Ip, Jp, Kp, d);
        bcentv[d] = ijkc_prev[d] + 0.5 + eb.bc(Ip, Jp, Kp, d);
      }
    } else {
      for (int dir = 0; dir < 3; ++dir) {
        if (ijkc[dir] != ijkc_prev[dir]) {
          const double face_sgn = (ijkc[dir] > ijkc_prev[dir]) ? -1. : 1.;
          normal[dir] = face_sgn;
          bcentv[dir] = static_cast<double>(ijkc[dir]) + 0.5 * face_sgn;
        } else {
          normal[dir] = 0.;
          bcentv[dir] = static_cast<double>(ijkc[dir]);
        }
      }
      const double len = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
      for (int d = 0; d < 3; ++d) normal[d] /= len;
    }
    return true;
  }
  return false;
}

}  // namespace oracle
