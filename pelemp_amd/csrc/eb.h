// eb.h -- embedded boundaries in the droplet update (FP64 branch kernel):
//   eb_interp           SprayInterpolation.H:395-632 (with get_interp_mapping,
//                       :201-393): trilinear weights through the cut-cell
//                       centroids by a Newton inversion of the trilinear map,
//                       or inverse-distance weights near covered cells, the
//                       EB or small cells, a droplet behind the EB moved to
//                       the boundary centroid of the closest cell;
//   eb_wall             WallFunctions.H:124-157, the EB part of check_wall;
//   eb_redirect_kernel  the deposits eb_interp sends to a neighbouring cell.
// EB data in AMReX's layout (include/pelemp_spray.h, pelemp_eb_view); cell
// indices in these functions are relative to the domain's lower corner.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pelemp_spray.h"
#include "pm_math.h"

namespace pelemp {

struct EbDev {
  const uint32_t* __restrict__ flags;
  const double* __restrict__ vfrac;
  const double* __restrict__ ccent;
  const double* __restrict__ bcent;
  const double* __restrict__ bnorm;
  int lo[3];  // box lower corner minus the domain's lower corner
  int n[3];
  long long cstride;

  PM_HDI bool has(int i, int j, int k) const {
    return i >= lo[0] && i < lo[0] + n[0] && j >= lo[1] && j < lo[1] + n[1] && k >= lo[2] && k < lo[2] + n[2];
  }
  PM_HDI long long off(int i, int j, int k) const {
    return (static_cast<long long>(k - lo[2]) * n[1] + (j - lo[1])) * n[0] + (i - lo[0]);
  }
  // outside the box: a regular cell, fully connected, volume fraction 1
  PM_HDI uint32_t flag(int i, int j, int k) const { return has(i, j, k) ? flags[off(i, j, k)] : 0x7FFFFFE0u; }
  PM_HDI bool regular(int i, int j, int k) const { return (flag(i, j, k) & 3u) == 0u; }
  PM_HDI bool single(int i, int j, int k) const { return (flag(i, j, k) & 3u) == 1u; }
  PM_HDI bool covered(int i, int j, int k) const { return (flag(i, j, k) & 3u) == 3u; }
  PM_HDI double vf(int i, int j, int k) const { return has(i, j, k) ? vfrac[off(i, j, k)] : 1.; }
  PM_HDI double cc(int i, int j, int k, int d) const { return has(i, j, k) ? ccent[off(i, j, k) + d * cstride] : 0.; }
  PM_HDI double bc(int i, int j, int k, int d) const { return has(i, j, k) ? bcent[off(i, j, k) + d * cstride] : 0.; }
  PM_HDI double bn(int i, int j, int k, int d) const { return has(i, j, k) ? bnorm[off(i, j, k) + d * cstride] : 0.; }
};

// the trilinear map's coefficients for component dir (SprayInterpolation.H:
// 207-318): x(xi,eta,zeta) - pos = a0 + a1 xi + a2 eta + a3 zeta + a4 xi eta
// + a5 xi zeta + a6 eta zeta + a7 xi eta zeta
struct EbMapCoef {
  double a1, a2, a3, a4, a5, a6, a7;
};

PM_HDI EbMapCoef eb_coef(int dir, const double (*nodes)[3]) {
  EbMapCoef c;
  c.a1 = (nodes[1][dir] - nodes[0][dir]);
  c.a2 = (nodes[2][dir] - nodes[0][dir]);
  c.a3 = (nodes[4][dir] - nodes[0][dir]);
  c.a4 = (nodes[0][dir] - nodes[1][dir] + nodes[3][dir] - nodes[2][dir]);
  c.a5 = (nodes[0][dir] - nodes[1][dir] - nodes[4][dir] + nodes[5][dir]);
  c.a6 = (nodes[0][dir] - nodes[2][dir] - nodes[4][dir] + nodes[6][dir]);
  c.a7 = (nodes[1][dir] - nodes[3][dir] + nodes[2][dir] + nodes[4][dir] - nodes[5][dir] + nodes[7][dir] -
          nodes[6][dir] - nodes[0][dir]);
  return c;
}

// get_interp_mapping: Newton with Cramer's rule, at most 10 steps, until
// the largest change is at most 1e-3
PM_HDI void eb_interp_mapping(const double* pos, const double (*nodes)[3], double& xi, double& eta, double& zeta) {
  EbMapCoef C[3];
  double a0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    C[d] = eb_coef(d, nodes);
    a0[d] = (nodes[0][d] - pos[d]);
  }
  int lc = 0;
  double err = 1.0;
  while (err > 1.0e-3 && lc < 10) {
    double f[3], dxi[3], deta[3], dzeta[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const EbMapCoef& c = C[d];
      f[d] = a0[d] + c.a1 * xi + c.a2 * eta + c.a3 * zeta + c.a4 * xi * eta + c.a5 * xi * zeta + c.a6 * eta * zeta +
             c.a7 * xi * eta * zeta;
      dxi[d] = c.a1 + c.a4 * eta + c.a5 * zeta + c.a7 * eta * zeta;
      deta[d] = c.a2 + c.a4 * xi + c.a6 * zeta + c.a7 * xi * zeta;
      dzeta[d] = c.a3 + c.a5 * xi + c.a6 * eta + c.a7 * xi * eta;
    }
    const double detJ = dxi[0] * (deta[1] * dzeta[2] - dzeta[1] * deta[2]) -
                        deta[0] * (dxi[1] * dzeta[2] - dzeta[1] * dxi[2]) +
                        dzeta[0] * (dxi[1] * deta[2] - deta[1] * dxi[2]);
    const double detJ_xi = f[0] * (deta[1] * dzeta[2] - dzeta[1] * deta[2]) - deta[0] * (f[1] * dzeta[2] - dzeta[1] * f[2]) +
                           dzeta[0] * (f[1] * deta[2] - deta[1] * f[2]);
    const double detJ_eta = dxi[0] * (f[1] * dzeta[2] - dzeta[1] * f[2]) - f[0] * (dxi[1] * dzeta[2] - dzeta[1] * dxi[2]) +
                            dzeta[0] * (dxi[1] * f[2] - f[1] * dxi[2]);
    const double detJ_zeta = dxi[0] * (deta[1] * f[2] - f[1] * deta[2]) - deta[0] * (dxi[1] * f[2] - f[1] * dxi[2]) +
                             f[0] * (dxi[1] * deta[2] - deta[1] * dxi[2]);
    const double new_xi = xi - detJ_xi / detJ;
    const double new_eta = eta - detJ_eta / detJ;
    const double new_zeta = zeta - detJ_zeta / detJ;
    lc += 1;
    const double e0 = fabs(xi - new_xi), e1 = fabs(eta - new_eta), e2 = fabs(zeta - new_zeta);
    err = e0 > e1 ? (e0 > e2 ? e0 : e2) : (e1 > e2 ? e1 : e2);
    xi = new_xi;
    eta = new_eta;
    zeta = new_zeta;
  }
}

// Result of eb_interp for one droplet: the 8 stencil cells (domain-relative)
// and weights in the reference's order
struct EbStencil {
  int idx[8][3];
  double w[8];
};

// Returns 0: all stencil cells regular (use the trilinear weights), 1: EB
// weights in st, -1: no acceptable cell in the stencil.  pos and ijkc may
// change (a droplet behind the EB, the closest acceptable cell).
PM_HDI int eb_interp(const EbDev& eb, double* pos, int* ijkc, const int* ijk, const double* plo,
                     const double* dx, const double* dxi, double min_eb_vfrac, EbStencil& st) {
  const int ip = ijkc[0], jp = ijkc[1], kp = ijkc[2];
  double normpos[3], normposp[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    normpos[d] = (pos[d] - plo[d]) * dxi[d];
    normposp[d] = normpos[d] - 0.5;
  }
  const int ic = ijk[0], jc = ijk[1], kc = ijk[2];
  bool do_fe_interp = false;
#pragma unroll
  for (int kk = kc - 1; kk <= kc; kk++)
#pragma unroll
    for (int jj = jc - 1; jj <= jc; jj++)
#pragma unroll
      for (int ii = ic - 1; ii <= ic; ii++)
        if (!eb.regular(ii, jj, kk)) do_fe_interp = true;
  if (!do_fe_interp) return 0;
  const double tolerance = 2.220446049250313e-16;
  double par_dot_EB = 2.;
  double cent_dot_EB = 1.;
  const uint32_t pflag = eb.flag(ip, jp, kp);
  if ((pflag & 3u) == 1u) {
    double normal[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) normal[d] = -eb.bn(ip, jp, kp, d);
    par_dot_EB = (normposp[0] - (ip + eb.bc(ip, jp, kp, 0))) * normal[0] +
                 (normposp[1] - (jp + eb.bc(ip, jp, kp, 1))) * normal[1] +
                 (normposp[2] - (kp + eb.bc(ip, jp, kp, 2))) * normal[2];
    cent_dot_EB = (eb.cc(ip, jp, kp, 0) - eb.bc(ip, jp, kp, 0)) * normal[0] +
                  (eb.cc(ip, jp, kp, 1) - eb.bc(ip, jp, kp, 1)) * normal[1] +
                  (eb.cc(ip, jp, kp, 2) - eb.bc(ip, jp, kp, 2)) * normal[2];
  }
  int covered = 0;
#pragma unroll
  for (int kk = kc - 1; kk <= kc; kk++)
#pragma unroll
    for (int jj = jc - 1; jj <= jc; jj++)
#pragma unroll
      for (int ii = ic - 1; ii <= ic; ii++) {
        const int bit = 5 + (ii - ip + 1) + 3 * (jj - jp + 1) + 9 * (kk - kp + 1);
        if (!((pflag >> bit) & 1u)) covered += 1;
      }
  const bool in_covered = (pflag & 3u) == 3u;
  const bool behind_EB = (par_dot_EB < tolerance);
  const bool bad_part = in_covered || behind_EB;
  const bool near_covered = (covered > 0);
  const bool near_EB = (par_dot_EB < cent_dot_EB);
  const bool in_small_cell = (eb.vf(ip, jp, kp) < min_eb_vfrac);
  if (near_covered || near_EB || in_small_cell || bad_part) {
    int aindx = 0;
    int closest_aindx = -1;
    double sum_wt = 0.;
    double min_dis = 1.E12;
    bool at_cent = false;
    for (int kk = kc - 1; kk <= kc; kk++) {
      for (int jj = jc - 1; jj <= jc; jj++) {
        const double y = static_cast<double>(jj);
        for (int ii = ic - 1; ii <= ic; ii++) {
          const double x = static_cast<double>(ii);
          st.idx[aindx][0] = ii;
          st.idx[aindx][1] = jj;
          st.idx[aindx][2] = kk;
          if (!eb.covered(ii, jj, kk) && eb.vf(ii, jj, kk) > min_eb_vfrac) {
            const double dlx = normposp[0] - (x + eb.cc(ii, jj, kk, 0));
            const double dly = normposp[1] - (y + eb.cc(ii, jj, kk, 1));
            const double dlz = normposp[2] - (static_cast<double>(kk) + eb.cc(ii, jj, kk, 2));
            const double len = sqrt(dlx * dlx + dly * dly + dlz * dlz);
            if (len == 0.) {
              st.w[aindx] = 1.;
              at_cent = true;
            } else {
              st.w[aindx] = 1. / len;
              sum_wt += 1. / len;
            }
            if (len < min_dis) {
              closest_aindx = aindx;
              min_dis = len;
            }
          } else {
            st.w[aindx] = 0.;
          }
          aindx++;
        }
      }
    }
    if (closest_aindx < 0) return -1;
    if (at_cent) {
      for (aindx = 1; aindx < 8; ++aindx) st.w[aindx] = 0.;
    } else {
      for (aindx = 0; aindx < 8; ++aindx) st.w[aindx] /= sum_wt;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) ijkc[d] = st.idx[closest_aindx][d];
    if (bad_part) {
#pragma unroll
      for (int dir = 0; dir < 3; ++dir)
        pos[dir] = (static_cast<double>(ijkc[dir]) + 0.5 + eb.bc(ijkc[0], ijkc[1], ijkc[2], dir)) * dx[dir] + plo[dir];
    }
  } else {
    double nodes[8][3];
    int lc = 0;
#pragma unroll
    for (int kk = kc - 1; kk <= kc; kk++)
#pragma unroll
      for (int jj = jc - 1; jj <= jc; jj++)
#pragma unroll
        for (int ii = ic - 1; ii <= ic; ii++) {
          nodes[lc][0] = (ii + 0.5 + eb.cc(ii, jj, kk, 0));
          nodes[lc][1] = (jj + 0.5 + eb.cc(ii, jj, kk, 1));
          nodes[lc][2] = (kk + 0.5 + eb.cc(ii, jj, kk, 2));
          lc += 1;
        }
    double xi = normpos[0] - nodes[0][0];
    double eta = normpos[1] - nodes[0][1];
    double zeta = normpos[2] - nodes[0][2];
    eb_interp_mapping(normpos, nodes, xi, eta, zeta);
    const int order[8][3] = {{-1, -1, -1}, {0, -1, -1}, {0, 0, -1}, {-1, 0, -1},
                             {-1, -1, 0},  {0, -1, 0},  {0, 0, 0},  {-1, 0, 0}};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      st.idx[a][0] = ic + order[a][0];
      st.idx[a][1] = jc + order[a][1];
      st.idx[a][2] = kc + order[a][2];
    }
    st.w[0] = 1. - xi - eta - zeta + xi * eta + xi * zeta + eta * zeta - xi * eta * zeta;
    st.w[1] = xi * (1. - eta - zeta + eta * zeta);
    st.w[2] = xi * eta - xi * eta * zeta;
    st.w[3] = eta * (1. - xi - zeta + xi * zeta);
    st.w[4] = zeta * (1. - xi - eta + xi * eta);
    st.w[5] = xi * zeta - xi * eta * zeta;
    st.w[6] = xi * eta * zeta;
    st.w[7] = eta * zeta - xi * eta * zeta;
  }
  return 1;
}

// check_wall's EB part (WallFunctions.H:124-157): ijkc the droplet's new
// cell, ijkc_prev the gather's cell; the wall's normal and a point on it
// (cell units).  Returns whether there is a wall to test.
PM_HDI bool eb_wall(const EbDev& eb, const int* ijkc, const int* ijkc_prev, double* normal, double* bcentv) {
  if (eb.single(ijkc[0], ijkc[1], ijkc[2])) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      normal[d] = -eb.bn(ijkc[0], ijkc[1], ijkc[2], d);
      bcentv[d] = ijkc[d] + 0.5 + eb.bc(ijkc[0], ijkc[1], ijkc[2], d);
    }
    return true;
  }
  if (eb.covered(ijkc[0], ijkc[1], ijkc[2])) {
    if (eb.single(ijkc_prev[0], ijkc_prev[1], ijkc_prev[2])) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        normal[d] = -eb.bn(ijkc_prev[0], ijkc_prev[1], ijkc_prev[2], d);
        bcentv[d] = ijkc_prev[d] + 0.5 + eb.bc(ijkc_prev[0], ijkc_prev[1], ijkc_prev[2], d);
      }
    } else {
#pragma unroll
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
      const double len = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
#pragma unroll
      for (int d = 0; d < 3; ++d) normal[d] /= len;
    }
    return true;
  }
  return false;
}

// a cell of the domain grown by one cell, as a linear index (x fastest)
PM_HDI uint32_t eb_cell_code(int x, int y, int z, const int* ncell) {
  return (static_cast<uint32_t>(z + 1) * (ncell[1] + 2) + static_cast<uint32_t>(y + 1)) * (ncell[0] + 2) +
         static_cast<uint32_t>(x + 1);
}

// Deposits that eb_interp sent to a neighbouring cell (the closest
// acceptable cell of the stencil): after the update pass, each cell next to
// the EB (domain or ghost cell, by eb_cell_code) adds those aimed at it, in
// sorted droplet order (neighbour cells by ascending key, each cell's
// droplets in order), after its own droplets' sum.  cstart: the pass's cell
// offsets (before the re-sort).
struct EbRedirectArgs {
  const int32_t* cells;
  int32_t ncells_list;
  const int32_t* cstart;
  const uint32_t* rkey;
  const double* rval;
  long long cap;
  int ncell[3], ntile[3], dom_lo[3];
  double* src;
  int src_lo[3], src_n[3];
  long long src_cstride;
  int comp[5 + PELEMP_MAX_FUEL];
  int on[5 + PELEMP_MAX_FUEL];
  int pelec;
  double scale[5 + PELEMP_MAX_FUEL];
  int is_ghost;
  int src_g;  // source ghost width: deposits beyond it are outside the source box (negative: no limit)
  int* err;
};

template <int NF>
__global__ void __launch_bounds__(128) eb_redirect_kernel(EbRedirectArgs r) {
  constexpr int NSRC = 5 + NF;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r.ncells_list) return;
  // target cell of the grown domain
  const uint32_t kc = static_cast<uint32_t>(r.cells[i]);
  const int gx0 = r.ncell[0] + 2, gy0 = r.ncell[1] + 2;
  const int cx = static_cast<int>(kc % gx0) - 1, cy = static_cast<int>((kc / gx0) % gy0) - 1;
  const int cz = static_cast<int>(kc / (gx0 * gy0)) - 1;
  auto key_of = [&](int x, int y, int z) {
    const uint32_t tile = (static_cast<uint32_t>(z / 4) * r.ntile[1] + (y / 4)) * r.ntile[0] + (x / 4);
    return tile * 64u + ((z % 4) * 4 + (y % 4)) * 4 + (x % 4);
  };
  uint32_t keys[27];
  int nk = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = cx + dx, y = cy + dy, z = cz + dz;
        if (x < 0 || x >= r.ncell[0] || y < 0 || y >= r.ncell[1] || z < 0 || z >= r.ncell[2]) continue;
        const uint32_t k = key_of(x, y, z);
        int j = nk++;
        while (j > 0 && keys[j - 1] > k) {
          keys[j] = keys[j - 1];
          --j;
        }
        keys[j] = k;
      }
  const int gx = r.dom_lo[0] + cx - r.src_lo[0], gy = r.dom_lo[1] + cy - r.src_lo[1], gz = r.dom_lo[2] + cz - r.src_lo[2];
  const bool in_box = r.src_g < 0 || (cx >= -r.src_g && cx < r.ncell[0] + r.src_g && cy >= -r.src_g &&
                                       cy < r.ncell[1] + r.src_g && cz >= -r.src_g && cz < r.ncell[2] + r.src_g);
  const bool in_src =
      in_box && gx >= 0 && gx < r.src_n[0] && gy >= 0 && gy < r.src_n[1] && gz >= 0 && gz < r.src_n[2];
  const long long cell = (static_cast<long long>(gz) * r.src_n[1] + gy) * r.src_n[0] + gx;
  double v[NSRC];
  bool any = false;
  for (int n = 0; n < nk; ++n) {
    const int32_t s0 = r.cstart[keys[n]], s1 = r.cstart[keys[n] + 1];
    for (int32_t s = s0; s < s1; ++s) {
      if (r.rkey[s] != kc) continue;
      if (!in_src) {
        if (!r.is_ghost) atomicCAS(r.err, 0, PELEMP_E_SRC_BOX);
        return;
      }
      if (!any) {
#pragma unroll
        for (int q = 0; q < NSRC; ++q) v[q] = r.on[q] ? r.src[r.comp[q] * r.src_cstride + cell] : 0.;
        any = true;
      }
#pragma unroll
      for (int q = 0; q < NSRC; ++q) {
        if (!r.on[q]) continue;
        const double c = r.rval[q * r.cap + s];
        v[q] = r.pelec ? v[q] + c * r.scale[q] : v[q] + c;
      }
    }
  }
  if (any) {
#pragma unroll
    for (int q = 0; q < NSRC; ++q)
      if (r.on[q]) r.src[r.comp[q] * r.src_cstride + cell] = v[q];
  }
}

}  // namespace pelemp
