// spray_host.hip -- C ABI of the spray droplet update (include/pelemp_spray.h).
//
// Owns the device-resident droplet set of one AMR level: hot droplet fields
// in ping-pong SoA buffers sorted by cell (tile-major key), cold per-slot
// data (id, breakup state when breakup is off), the sort permutation and the
// CSR cell offsets.  One update pass = fused update kernel + deterministic
// counting sort; no host synchronisation inside the pass chain.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <fstream>
#include <limits>
#include <sstream>
#include <cmath>
#include <cstring>
#include <string>

#include "droplet_physics.h"
#include "gas_props.h"
#include "pm_math.h"
#include "sort_kernels.h"
#include "spray_ctx.h"
#include "update_kernel.h"
#include "branch_kernels.h"
#include "derive_kernels.h"
#include "transfer_kernels.h"
#include "units.h"
#include "inject_kernels.h"
#include "particle_io.h"

struct pelemp_spray_ctx {
  pelemp::Ctx c;
};

namespace pelemp {

thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

#define PM_HIP(expr)                                                           \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
      return PELEMP_E_HIP;                                                     \
    }                                                                          \
  } while (0)

hipEvent_t ctx_event(Ctx& c) {
  if (!c.ev_pool.empty()) {
    hipEvent_t e = c.ev_pool.back();
    c.ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void ctx_time_begin(Ctx& c, int slot, hipStream_t s, hipEvent_t* ev) {
  (void)slot;
  *ev = nullptr;
  if (!c.timing) return;
  *ev = ctx_event(c);
  if (*ev) hipEventRecord(*ev, s);
}

void ctx_time_end(Ctx& c, int slot, hipStream_t s, hipEvent_t ev0) {
  if (!c.timing || !ev0) return;
  hipEvent_t ev1 = ctx_event(c);
  if (!ev1) return;
  hipEventRecord(ev1, s);
  c.ev_pending.push_back({slot, {ev0, ev1}});
}

static void harvest_timing(Ctx& c) {
  for (auto& pe : c.ev_pending) {
    hipEventSynchronize(pe.second.second);
    float ms = 0.f;
    hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
    c.t_ms[pe.first] += ms;
    c.t_n[pe.first] += 1;
    c.ev_pool.push_back(pe.second.first);
    c.ev_pool.push_back(pe.second.second);
  }
  c.ev_pending.clear();
}

// ---------------------------------------------------------------------------
// helper kernels

// AoS host-layout particles -> SoA buffer `cur` (storage order = input order)
template <class R, bool POS64>
__global__ void __launch_bounds__(256) ingest_kernel(int64_t n, int nf, int nsr, const double* __restrict__ pos,
                                                     const double* __restrict__ rdata,
                                                     const int64_t* __restrict__ id, R* __restrict__ f,
                                                     int64_t cap, int hot_brk, int32_t* __restrict__ slot,
                                                     int64_t slot_base, int64_t* __restrict__ cold_id,
                                                     double* __restrict__ cold_n0, double* __restrict__ cold_bm1,
                                                     double* __restrict__ cold_bm2, uint32_t* __restrict__ key,
                                                     double plo0, double plo1, double plo2, double dxi0,
                                                     double dxi1, double dxi2, int nc0, int nc1, int nc2,
                                                     int nt0, int nt1, int nt2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double plo[3] = {plo0, plo1, plo2};
  const double dxi[3] = {dxi0, dxi1, dxi2};
  const int nc[3] = {nc0, nc1, nc2};
  const int nt[3] = {nt0, nt1, nt2};
  int c[3];
  for (int q = 0; q < 3; ++q) {
    const double p = pos[3 * i + q];
    const double lxc = (p - plo[q]) * dxi[q];
    int ic = static_cast<int>(floor(lxc));
    ic = ic < 0 ? 0 : (ic >= nc[q] ? nc[q] - 1 : ic);
    c[q] = ic;
    if constexpr (POS64) {
      f[q * cap + i] = static_cast<R>(p);
    } else {
      double xi = lxc - static_cast<double>(ic);
      xi = xi < 0. ? 0. : xi;
      R xr = static_cast<R>(xi);
      if (xr >= R(1)) xr = R(0.99999994f);
      f[q * cap + i] = xr;
    }
  }
  const double* r = rdata + static_cast<int64_t>(nsr) * i;
  // Fields<NF>: VEL 3..5, T 6, DIA 7, Y 8..7+nf, ND 8+nf, FILM 9+nf, BM1 10+nf, BM2 11+nf, N0 12+nf
  for (int q = 0; q < 3; ++q) f[(3 + q) * cap + i] = static_cast<R>(r[q]);
  f[6 * cap + i] = static_cast<R>(r[3]);
  f[7 * cap + i] = static_cast<R>(r[4]);
  for (int s = 0; s < nf; ++s) f[(8 + s) * cap + i] = static_cast<R>(r[5 + s]);
  f[(8 + nf) * cap + i] = static_cast<R>(r[5 + nf]);  // numDens
  f[(9 + nf) * cap + i] = static_cast<R>(r[9 + nf]);  // FilmHght
  const int64_t sl = slot_base + i;
  if (hot_brk) {
    f[(10 + nf) * cap + i] = static_cast<R>(r[7 + nf]);
    f[(11 + nf) * cap + i] = static_cast<R>(r[8 + nf]);
    f[(12 + nf) * cap + i] = static_cast<R>(r[6 + nf]);
  } else {
    cold_n0[sl] = r[6 + nf];
    cold_bm1[sl] = r[7 + nf];
    cold_bm2[sl] = r[8 + nf];
  }
  slot[i] = static_cast<int32_t>(sl);
  cold_id[sl] = id[i];
  key[i] = tile_key(c[0], c[1], c[2], nt);
}

__device__ inline void decode_key(uint32_t k, const int* nt, int* c) {
  const uint32_t tile = k / kTileCells;
  const uint32_t l = k % kTileCells;
  const uint32_t txy = static_cast<uint32_t>(nt[0]) * nt[1];
  c[2] = static_cast<int>((tile / txy) * kTile + l / (kTile * kTile));
  c[1] = static_cast<int>(((tile % txy) / nt[0]) * kTile + (l / kTile) % kTile);
  c[0] = static_cast<int>((tile % nt[0]) * kTile + l % kTile);
}

// sorted view -> host-layout AoS
template <class R, bool POS64>
__global__ void __launch_bounds__(256) gather_kernel(const int32_t* __restrict__ total, int nf, int nsr,
                                                     const int32_t* __restrict__ perm, const R* __restrict__ f,
                                                     int64_t cap, int hot_brk, const int32_t* __restrict__ slot,
                                                     const int64_t* __restrict__ cold_id,
                                                     const double* __restrict__ cold_n0,
                                                     const double* __restrict__ cold_bm1,
                                                     const double* __restrict__ cold_bm2,
                                                     const uint32_t* __restrict__ key, double* __restrict__ pos,
                                                     double* __restrict__ rdata, int64_t* __restrict__ id,
                                                     double plo0, double plo1, double plo2, double dx0, double dx1,
                                                     double dx2, int nt0, int nt1, int nt2) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= *total) return;
  const int32_t o = perm[s];
  if constexpr (POS64) {
    for (int q = 0; q < 3; ++q) pos[3 * s + q] = static_cast<double>(f[q * cap + o]);
  } else {
    const int nt[3] = {nt0, nt1, nt2};
    int c[3];
    decode_key(key[o], nt, c);
    const double plo[3] = {plo0, plo1, plo2};
    const double dx[3] = {dx0, dx1, dx2};
    for (int q = 0; q < 3; ++q)
      pos[3 * s + q] = plo[q] + (static_cast<double>(c[q]) + static_cast<double>(f[q * cap + o])) * dx[q];
  }
  double* r = rdata + static_cast<int64_t>(nsr) * s;
  for (int q = 0; q < 3; ++q) r[q] = static_cast<double>(f[(3 + q) * cap + o]);
  r[3] = static_cast<double>(f[6 * cap + o]);
  r[4] = static_cast<double>(f[7 * cap + o]);
  for (int sp = 0; sp < nf; ++sp) r[5 + sp] = static_cast<double>(f[(8 + sp) * cap + o]);
  r[5 + nf] = static_cast<double>(f[(8 + nf) * cap + o]);
  r[9 + nf] = static_cast<double>(f[(9 + nf) * cap + o]);
  const int32_t sl = slot[o];
  if (hot_brk) {
    r[7 + nf] = static_cast<double>(f[(10 + nf) * cap + o]);
    r[8 + nf] = static_cast<double>(f[(11 + nf) * cap + o]);
    r[6 + nf] = static_cast<double>(f[(12 + nf) * cap + o]);
  } else {
    r[6 + nf] = cold_n0[sl];
    r[7 + nf] = cold_bm1[sl];
    r[8 + nf] = cold_bm2[sl];
  }
  id[s] = cold_id[sl];
}

// FP32 form: host-layout particles -> Fields32 + species masses.  `mass` is
// the droplet mass pi/6 rho_l(T) d^3 computed on the host.
__global__ void __launch_bounds__(256) ingest32_kernel(int64_t n, int nf, int nsr, const double* __restrict__ pos,
                                                       const double* __restrict__ rdata,
                                                       const double* __restrict__ mass,
                                                       const int64_t* __restrict__ id, float* __restrict__ f,
                                                       double* __restrict__ mk, int64_t cap,
                                                       int32_t* __restrict__ slot, int64_t slot_base,
                                                       int64_t* __restrict__ cold_id, double* __restrict__ cold_n0,
                                                       double* __restrict__ cold_bm1, double* __restrict__ cold_bm2,
                                                       uint32_t* __restrict__ key, double plo0, double plo1,
                                                       double plo2, double dxi0, double dxi1, double dxi2, int nc0,
                                                       int nc1, int nc2, int nt0, int nt1, int nt2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  using G = Fields32;
  const double plo[3] = {plo0, plo1, plo2};
  const double dxi[3] = {dxi0, dxi1, dxi2};
  const int nc[3] = {nc0, nc1, nc2};
  const int nt[3] = {nt0, nt1, nt2};
  int c[3];
  for (int q = 0; q < 3; ++q) {
    const double lxc = (pos[3 * i + q] - plo[q]) * dxi[q];
    int ic = static_cast<int>(floor(lxc));
    ic = ic < 0 ? 0 : (ic >= nc[q] ? nc[q] - 1 : ic);
    c[q] = ic;
    double xi = lxc - static_cast<double>(ic);
    xi = xi < 0. ? 0. : xi;
    float xr = static_cast<float>(xi);
    if (xr >= 1.f) xr = 0.99999994f;
    f[(G::POS + q) * cap + i] = xr;
  }
  const double* r = rdata + static_cast<int64_t>(nsr) * i;
  for (int q = 0; q < 3; ++q) f[(G::VEL + q) * cap + i] = static_cast<float>(r[q]);
  f[G::T * cap + i] = static_cast<float>(r[3]);
  f[G::DIA * cap + i] = static_cast<float>(r[4]);
  f[G::ND * cap + i] = static_cast<float>(r[5 + nf]);
  for (int s = 0; s < nf; ++s) mk[s * cap + i] = r[5 + s] * mass[i];
  const int64_t sl = slot_base + i;
  cold_n0[sl] = r[6 + nf];
  cold_bm1[sl] = r[7 + nf];
  cold_bm2[sl] = r[8 + nf];
  slot[i] = static_cast<int32_t>(sl);
  cold_id[sl] = id[i];
  key[i] = tile_key(c[0], c[1], c[2], nt);
}

__global__ void __launch_bounds__(256) gather32_kernel(const int32_t* __restrict__ total, int nf, int nsr,
                                                       const int32_t* __restrict__ perm, const float* __restrict__ f,
                                                       const double* __restrict__ mk, int64_t cap,
                                                       const int32_t* __restrict__ slot,
                                                       const int64_t* __restrict__ cold_id,
                                                       const double* __restrict__ cold_n0,
                                                       const double* __restrict__ cold_bm1,
                                                       const double* __restrict__ cold_bm2,
                                                       const uint32_t* __restrict__ key, double* __restrict__ pos,
                                                       double* __restrict__ rdata, int64_t* __restrict__ id,
                                                       double plo0, double plo1, double plo2, double dx0, double dx1,
                                                       double dx2, int nt0, int nt1, int nt2) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= *total) return;
  using G = Fields32;
  const int32_t o = perm[s];
  const int nt[3] = {nt0, nt1, nt2};
  int c[3];
  decode_key(key[o], nt, c);
  const double plo[3] = {plo0, plo1, plo2};
  const double dx[3] = {dx0, dx1, dx2};
  for (int q = 0; q < 3; ++q)
    pos[3 * s + q] = plo[q] + (static_cast<double>(c[q]) + static_cast<double>(f[(G::POS + q) * cap + o])) * dx[q];
  double* r = rdata + static_cast<int64_t>(nsr) * s;
  for (int q = 0; q < 3; ++q) r[q] = static_cast<double>(f[(G::VEL + q) * cap + o]);
  r[3] = static_cast<double>(f[G::T * cap + o]);
  r[4] = static_cast<double>(f[G::DIA * cap + o]);
  double m = 0.;
  for (int sp = 0; sp < nf; ++sp) m += mk[sp * cap + o];
  for (int sp = 0; sp < nf; ++sp) r[5 + sp] = (m > 0.) ? mk[sp * cap + o] / m : 0.;
  r[5 + nf] = static_cast<double>(f[G::ND * cap + o]);
  r[9 + nf] = 0.;
  const int32_t sl = slot[o];
  r[6 + nf] = cold_n0[sl];
  r[7 + nf] = cold_bm1[sl];
  r[8 + nf] = cold_bm2[sl];
  id[s] = cold_id[sl];
}

// particle-CFL time step: min over droplets of cfl / max_d(|u_d|/dx_d)
// (SprayParticles.cpp:112-124).  Positive doubles order like their bit
// patterns, so an integer atomicMin gives the exact minimum.
template <class R>
__global__ void __launch_bounds__(256) est_dt_kernel(const int32_t* __restrict__ total,
                                                     const int32_t* __restrict__ perm, const R* __restrict__ f,
                                                     int64_t cap, double dxi0, double dxi1, double dxi2, double cfl,
                                                     unsigned long long* __restrict__ out) {
  __shared__ double red[256];
  double m = 1.0e300;
  const int32_t n = *total;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int32_t o = perm[s];
    const double u0 = fabs(static_cast<double>(f[3 * cap + o])) * dxi0;
    const double u1 = fabs(static_cast<double>(f[4 * cap + o])) * dxi1;
    const double u2 = fabs(static_cast<double>(f[5 * cap + o])) * dxi2;
    double mx = u0;
    mx = (mx < u1) ? u1 : mx;
    mx = (mx < u2) ? u2 : mx;
    const double dtp = (mx > 0.) ? (cfl / mx) : 1.E50;
    m = dtp < m ? dtp : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(w)) {
      const double o2 = red[threadIdx.x + w];
      red[threadIdx.x] = o2 < red[threadIdx.x] ? o2 : red[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    unsigned long long bits;
    const double v = red[0];
    memcpy(&bits, &v, sizeof(bits));
    atomicMin(out, bits);
  }
}

__global__ void __launch_bounds__(256) math_eval_kernel(int fn, int64_t n, const double* __restrict__ x,
                                                        const double* __restrict__ y, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i];
  const double b = y ? y[i] : 0.0;
  double r = 0.0;
  switch (fn) {
    case 0: r = pm::exp(a); break;
    case 1: r = pm::log(a); break;
    case 2: r = pm::log1p(a); break;
    case 3: r = pm::pow(a, b); break;
    case 4: r = pm::cbrt(a); break;
    case 5: r = pm::sin(a); break;
    case 6: r = pm::cos(a); break;
    case 7: r = pm::atan2(a, b); break;
    case 8: r = pm::acos(a); break;
    case 9: r = sqrt(a); break;
    case 10: r = a / b; break;
    default: r = 0.0;
  }
  out[i] = r;
}

// ---------------------------------------------------------------------------

// Full stable sort of buffer cur by cell key.  keep_entries: the device
// count n_entries of storage entries is already set (droplets appended);
// otherwise it is the active count of the previous sort.
static int run_sort(Ctx& c, hipStream_t s, bool keep_entries = false) {
  hipEvent_t ev0;
  ctx_time_begin(c, 1, s, &ev0);
  if (!keep_entries)
    PM_HIP(hipMemcpyAsync(c.n_entries, c.cstart + c.ncells, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  PM_HIP(hipMemsetAsync(c.cnt, 0, sizeof(int32_t) * c.ncells, s));
  const int grid_s = 2048;
  hipLaunchKernelGGL(count_keys_kernel, dim3(grid_s), dim3(256), 0, s, c.key, c.n_entries, c.cnt, c.arrival);
  const int nb = (c.ncells + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(kScanThreads), 0, s, c.cnt, nullptr, nullptr, c.ncells,
                     c.cstart, c.block_sum, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(kScanThreads), 0, s, c.cstart, c.ncells, c.block_sum, nb, c.total,
                     c.n_long);
  hipLaunchKernelGGL(scatter_perm_kernel, dim3(grid_s), dim3(256), 0, s, c.key, c.n_entries, c.arrival, c.cstart,
                     c.perm, c.sp.counters);
  hipLaunchKernelGGL(order_cells_kernel, dim3((c.ncells + 255) / 256), dim3(256), 0, s, c.cstart, c.ncells,
                     c.perm, c.long_list, c.n_long, c.mov_lo);
  c.counters_dirty = false;
  // arrival is free once the scatter has read it: the merge buffer
  hipLaunchKernelGGL(sort_long_groups_kernel, dim3(kLongSortGrid), dim3(kLongSortThreads), 0, s, c.long_list,
                     c.n_long, c.perm, c.arrival);
  hipLaunchKernelGGL(split_tiles_kernel, dim3((c.ntiles + 255) / 256), dim3(256), 0, s, c.cstart, c.ntiles, c.sp);
  PM_HIP(hipGetLastError());
  ctx_time_end(c, 1, s, ev0);
  return PELEMP_OK;
}

// Film thickness per cell from the film droplets of buffer cur.
static int run_film_fab(Ctx& c, hipStream_t s) {
  const double face_area = c.geom.dx[0] * c.geom.dx[1] * c.geom.dx[2];  // SprayParticles.cpp:281
  const dim3 grid(static_cast<unsigned>((c.ncells + 255) / 256)), block(256);
  auto run = [&](auto nft) {
    constexpr int NF = decltype(nft)::value;
    hipLaunchKernelGGL((film_fab_kernel<NF>), grid, block, 0, s, c.cstart, c.perm,
                       static_cast<const double*>(c.fbuf[c.cur]), c.cap, c.ncell[0], c.ncell[1], c.ncell[2],
                       c.ntile[0], c.ntile[1], face_area, c.wf);
  };
  switch (c.nf) {
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 6: run(std::integral_constant<int, 6>{}); break;
    default: set_last_error("unsupported fuel count"); return PELEMP_E_BAD_CONFIG;
  }
  PM_HIP(hipGetLastError());
  return PELEMP_OK;
}

// Secondary droplets of the records the update pass wrote into buffer
// 1 - cur: counts, exclusive scan, creation, counters.
static int run_children(Ctx& c, const PassParams& p, hipStream_t s) {
  const int32_t* n_ptr = c.cstart + c.ncells;  // entries written by the update pass
  int32_t* count = c.sb_count;
  int64_t* offset = c.sb_offset;
  const unsigned nb = static_cast<unsigned>((c.cap + kScanBlock - 1) / kScanBlock);
  SBGeom gm;
  for (int d = 0; d < 3; ++d) {
    gm.plo[d] = c.geom.plo[d];
    gm.phi[d] = c.geom.phi[d];
    gm.dxi[d] = 1. / c.geom.dx[d];
    gm.bndry_lo[d] = c.geom.bndry_lo[d];
    gm.ncell[d] = c.ncell[d];
    gm.ntile[d] = c.ntile[d];
  }
  auto run = [&](auto tag, auto nft) {
    using M = decltype(tag);
    constexpr int NF = decltype(nft)::value;
    hipLaunchKernelGGL((sb_count_kernel<NF>), dim3(2048), dim3(256), 0, s, n_ptr, c.sb_type, c.sb_data, c.cap,
                       c.params.breakup_parcel_factor, count);
    hipLaunchKernelGGL(sb_scan_block_kernel, dim3(nb), dim3(kScanBlock), 0, s, count, offset, n_ptr, c.sb_block);
    hipLaunchKernelGGL(sb_scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, s, c.sb_block, n_ptr);
    hipLaunchKernelGGL(sb_scan_add_kernel, dim3(nb), dim3(kScanBlock), 0, s, offset, n_ptr, c.sb_block);
    hipLaunchKernelGGL(sb_total_kernel, dim3(1), dim3(64), 0, s, n_ptr, offset, count, c.sb_ids, c.n_new, c.cap);
    // parent record of every child: mark first children, inclusive max-scan
    hipLaunchKernelGGL(sb_parent_init_kernel, dim3(2048), dim3(256), 0, s, c.sb_ids, c.cap, c.sb_parent);
    hipLaunchKernelGGL(sb_parent_mark_kernel, dim3(2048), dim3(256), 0, s, n_ptr, count, offset, c.cap,
                       c.sb_parent);
    hipLaunchKernelGGL(sb_maxscan_block_kernel, dim3(nb), dim3(kScanBlock), 0, s, c.sb_parent, c.sb_ids, c.cap,
                       c.sb_pblock);
    hipLaunchKernelGGL(sb_maxscan_sums_kernel, dim3(1), dim3(kScanBlock), 0, s, c.sb_pblock, c.sb_ids, c.cap);
    hipLaunchKernelGGL(sb_maxscan_add_kernel, dim3(nb), dim3(kScanBlock), 0, s, c.sb_parent, c.sb_ids, c.cap,
                       c.sb_pblock);
    hipLaunchKernelGGL((sb_create_kernel<NF>), dim3(4096), dim3(256), 0, s, n_ptr, c.sb_type, c.sb_data, offset,
                       c.sb_parent, c.sb_ids, make_fuel_const<double, NF>(c.sd, c.mw.data()), p.sub_dt,
                       c.params.breakup_parcel_factor, c.params.rng_seed, static_cast<double*>(c.fbuf[1 - c.cur]),
                       c.cap, c.slot[1 - c.cur], c.cold_id, c.slot_cap, c.key, c.arrival, c.mov_hi, gm, c.err);
    hipLaunchKernelGGL(sb_advance_kernel, dim3(1), dim3(64), 0, s, c.sb_ids);
    (void)sizeof(M);
  };
  if (c.mech == PELEMP_MECH_DODECANE_AIR && c.nf == 1) {
    run(Mech_dodecane_air{}, std::integral_constant<int, 1>{});
  } else if (c.mech == PELEMP_MECH_DECANE_DODECANE_AIR && c.nf == 2) {
    run(Mech_decane_dodecane_air{}, std::integral_constant<int, 2>{});
  } else if (c.mech == PELEMP_MECH_ALKANE6_AIR && c.nf == 6) {
    run(Mech_alkane6_air{}, std::integral_constant<int, 6>{});
  } else if (c.mech == PELEMP_MECH_DECANE_AIR && c.nf == 1) {
    run(Mech_decane_air{}, std::integral_constant<int, 1>{});
  } else {
    set_last_error("unsupported mechanism / fuel count combination");
    return PELEMP_E_BAD_CONFIG;
  }
  PM_HIP(hipGetLastError());
  return PELEMP_OK;
}

// Deposits sent by eb_interp to a neighbouring cell, added by the cells next
// to the EB after the update pass (before the re-sort rewrites cstart).
static int run_eb_redirect(Ctx& c, const PassParams& p, hipStream_t s) {
  EbRedirectArgs r{};
  r.cells = c.eb_cells;
  r.ncells_list = c.n_eb_cells;
  r.cstart = c.cstart;
  r.rkey = c.eb_rkey;
  r.rval = c.eb_rval;
  r.cap = c.cap;
  const pelemp_fab_view& sv = *p.source;
  for (int d = 0; d < 3; ++d) {
    r.ncell[d] = c.ncell[d];
    r.ntile[d] = c.ntile[d];
    r.dom_lo[d] = c.geom.domain_lo[d];
    r.src_lo[d] = sv.lo[d];
    r.src_n[d] = sv.hi[d] - sv.lo[d] + 1;
  }
  r.src = sv.data;
  r.src_cstride = static_cast<long long>(r.src_n[0]) * r.src_n[1] * r.src_n[2];
  const int nsrc = 5 + c.nf;
  for (int q = 0; q < nsrc; ++q) {
    r.comp[q] = (q < 3) ? c.comps.momSrcIndx + q
                        : (q == 3) ? c.comps.rhoSrcIndx
                                   : (q < 4 + c.nf) ? c.comps.specSrcIndx + (q - 4) : c.comps.engSrcIndx;
    r.on[q] = (q < 3) ? (c.sd.mom_trans != 0) : (q < 4 + c.nf) ? (c.sd.mass_trans != 0) : 1;
    r.scale[q] = c.src_scale[q];
  }
  r.pelec = c.pelec ? 1 : 0;
  r.is_ghost = p.is_ghost;
  r.src_g = p.source_ghosts;
  r.err = c.err;
  const dim3 grid(static_cast<unsigned>((c.n_eb_cells + 127) / 128)), block(128);
  switch (c.nf) {
    case 1: hipLaunchKernelGGL(eb_redirect_kernel<1>, grid, block, 0, s, r); break;
    case 2: hipLaunchKernelGGL(eb_redirect_kernel<2>, grid, block, 0, s, r); break;
    case 6: hipLaunchKernelGGL(eb_redirect_kernel<6>, grid, block, 0, s, r); break;
    default: set_last_error("unsupported fuel count"); return PELEMP_E_BAD_CONFIG;
  }
  PM_HIP(hipGetLastError());
  return PELEMP_OK;
}

// Re-sort after an update pass (incremental form, see sort_kernels.h).
// n_src: device count of entries in the new buffer when it differs from the
// previous sort's count (secondary droplets appended).
// log_slot: where the droplet count at the pass start is logged (or null).
static int run_sort_incremental(Ctx& c, hipStream_t s, const int32_t* n_src, int32_t* log_slot) {
  hipEvent_t ev0;
  ctx_time_begin(c, 1, s, &ev0);
  const unsigned gc = static_cast<unsigned>((c.ncells + 255) / 256);
  const int nb = (c.ncells + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(kScanThreads), 0, s, c.mov_lo, c.stay_cnt, c.mov_hi, c.ncells,
                     c.cstart, c.block_sum, n_src, c.n_entries, log_slot);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(kScanThreads), 0, s, c.cstart, c.ncells, c.block_sum, nb, c.total,
                     c.n_long);
  hipLaunchKernelGGL(scatter_incr_kernel, dim3(2048), dim3(256), 0, s, c.key, c.n_entries, c.arrival, c.cstart,
                     c.mov_lo, c.stay_cnt, c.perm, c.sp);
  hipLaunchKernelGGL(order_movers_kernel, dim3(gc), dim3(256), 0, s, c.cstart, c.mov_lo, c.stay_cnt, c.mov_hi,
                     c.ncells, c.perm, c.long_list, c.n_long, c.ntiles, c.sp);
  c.counters_dirty = false;
  hipLaunchKernelGGL(sort_long_groups_kernel, dim3(kLongSortGrid), dim3(kLongSortThreads), 0, s, c.long_list,
                     c.n_long, c.perm, c.arrival);
  PM_HIP(hipGetLastError());
  ctx_time_end(c, 1, s, ev0);
  return PELEMP_OK;
}

static int alloc(Ctx& c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (hipMalloc(p, bytes) != hipSuccess) {
    set_last_error("hipMalloc failed (" + std::to_string(bytes) + " bytes)");
    return PELEMP_E_NOMEM;
  }
  c.bytes += bytes;
  return PELEMP_OK;
}

// Buffers of the branch kernel: film thickness per cell, secondary-droplet
// records and their scans, the new-droplet counters.
static int alloc_branch_buffers(Ctx& c) {
  int st = PELEMP_OK;
  st |= alloc(c, reinterpret_cast<void**>(&c.wf), 8 * static_cast<size_t>(c.ncells));
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_type), 4 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_data), 8 * static_cast<size_t>(kSBFields + c.nf) * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_count), 4 * static_cast<size_t>(c.cap));
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_offset), 8 * static_cast<size_t>(c.cap));
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_block), 8 * (static_cast<size_t>(c.cap) / kScanBlock + 2));
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_parent), 4 * static_cast<size_t>(c.cap));
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_pblock), 4 * (static_cast<size_t>(c.cap) / kScanBlock + 2));
  st |= alloc(c, reinterpret_cast<void**>(&c.sb_ids), 8 * 4);
  st |= alloc(c, reinterpret_cast<void**>(&c.n_new), 64);
  return st;
}

static void free_ctx(Ctx& c) {
  void* ptrs[] = {c.fbuf[0], c.fbuf[1], c.mbuf[0], c.mbuf[1], c.slot[0], c.slot[1], c.perm, c.cstart, c.cnt,
                  c.arrival, c.long_list, c.mov_lo, c.key,
                  c.block_sum, c.scalars, c.dscratch, c.cold_id, c.cold_n0, c.cold_bm1, c.cold_bm2,
                  c.count_log, c.surr_w, c.surr_cnt, c.wf, c.sb_type, c.sb_data, c.sb_count, c.sb_offset,
                  c.sb_block, c.sb_parent, c.sb_pblock, c.sb_ids, c.n_new, c.eb_cells, c.eb_rkey, c.eb_rval,
                  c.pelec_conv, c.sp.tile_slot, c.sp.items, c.sp.list, c.sp.part, c.sp.seg_stay};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& pe : c.ev_pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (hipEvent_t e : c.ev_pool) (void)hipEventDestroy(e);
}

int mech_ns(int mech) {
  switch (mech) {
    case PELEMP_MECH_DODECANE_AIR: return Mech_dodecane_air::NS;
    case PELEMP_MECH_DECANE_DODECANE_AIR: return Mech_decane_dodecane_air::NS;
    case PELEMP_MECH_ALKANE6_AIR: return Mech_alkane6_air::NS;
    case PELEMP_MECH_DECANE_AIR: return Mech_decane_air::NS;
    default: return -1;
  }
}

void mech_mw(int mech, std::vector<double>& mw) {
  mw.clear();
  auto fill = [&](auto tag) {
    using M = decltype(tag);
    for (int n = 0; n < M::NS; ++n) mw.push_back(M::W[n]);
  };
  switch (mech) {
    case PELEMP_MECH_DODECANE_AIR: fill(Mech_dodecane_air{}); break;
    case PELEMP_MECH_DECANE_DODECANE_AIR: fill(Mech_decane_dodecane_air{}); break;
    case PELEMP_MECH_ALKANE6_AIR: fill(Mech_alkane6_air{}); break;
    case PELEMP_MECH_DECANE_AIR: fill(Mech_decane_air{}); break;
    default: break;
  }
}

void mech_h(int mech, double T, double* h) {
  auto run = [&](auto tag) {
    using M = decltype(tag);
    GasProps<M>::T2Hi(T, h);
  };
  switch (mech) {
    case PELEMP_MECH_DODECANE_AIR: run(Mech_dodecane_air{}); break;
    case PELEMP_MECH_DECANE_DODECANE_AIR: run(Mech_decane_dodecane_air{}); break;
    case PELEMP_MECH_ALKANE6_AIR: run(Mech_alkane6_air{}); break;
    case PELEMP_MECH_DECANE_AIR: run(Mech_decane_air{}); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// Appending droplets (add_particles, injection, lattice initialisation)

struct RhoCoef {
  double a[4 * PELEMP_MAX_FUEL];
};

// FP32 form: droplet mass pi/6 rho_l(T) d^3 with the mixture density
// 1/sum(Y/rhoL(T)), the same expression the host uses for set_particles
__global__ void __launch_bounds__(256) drop_mass_kernel(int64_t n, int nf, int nsr, const double* __restrict__ rdata,
                                                        RhoCoef rc, double* __restrict__ mass) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* r = rdata + static_cast<int64_t>(nsr) * i;
  const double T = r[3];
  double inv = 0.;
  for (int s = 0; s < nf; ++s) {
    const double* a = &rc.a[4 * s];
    inv += r[5 + s] / (a[0] + T * (a[1] + T * (a[2] + T * a[3])));
  }
  mass[i] = kPIc / 6. * (1. / inv) * (r[4] * r[4] * r[4]);
}

// next id and next free slot
static int read_counters(Ctx& c, int64_t* next_id, int64_t* next_slot) {
  if (c.branches) {
    int64_t ids[4];
    PM_HIP(hipMemcpy(ids, c.sb_ids, sizeof(ids), hipMemcpyDeviceToHost));
    *next_id = ids[0];
    *next_slot = ids[1];
  } else {
    *next_id = c.next_id;
    *next_slot = c.nslots;
  }
  return PELEMP_OK;
}

static int write_counters(Ctx& c, int64_t next_id, int64_t next_slot) {
  c.next_id = next_id;
  c.nslots = next_slot;
  if (c.branches) {
    const int64_t ids[2] = {next_id, next_slot};
    PM_HIP(hipMemcpy(c.sb_ids, ids, sizeof(ids), hipMemcpyHostToDevice));
  }
  return PELEMP_OK;
}

// Give the live droplets the slots 0..total-1 (their sorted positions).
static int compact_slots(Ctx& c, int64_t* next_slot) {
  int32_t total = 0;
  PM_HIP(hipMemcpy(&total, c.cstart + c.ncells, 4, hipMemcpyDeviceToHost));
  const bool with_bm = !c.hot_brk;
  const size_t n = std::max<int32_t>(total, 1);
  int64_t* tid = nullptr;
  double* tbm = nullptr;
  PM_HIP(hipMalloc(&tid, 8 * n));
  if (with_bm && hipMalloc(&tbm, 3 * 8 * n) != hipSuccess) {
    (void)hipFree(tid);
    set_last_error("hipMalloc failed (slot compaction)");
    return PELEMP_E_NOMEM;
  }
  hipLaunchKernelGGL(slot_compact_kernel, dim3(1024), dim3(256), 0, 0, c.cstart + c.ncells, c.perm, c.slot[c.cur],
                     c.cold_id, c.cold_n0, c.cold_bm1, c.cold_bm2, with_bm ? 1 : 0, tid, tbm, tbm ? tbm + n : nullptr,
                     tbm ? tbm + 2 * n : nullptr);
  PM_HIP(hipGetLastError());
  if (total > 0) {
    PM_HIP(hipMemcpy(c.cold_id, tid, 8 * static_cast<size_t>(total), hipMemcpyDeviceToDevice));
    if (with_bm) {
      PM_HIP(hipMemcpy(c.cold_n0, tbm, 8 * static_cast<size_t>(total), hipMemcpyDeviceToDevice));
      PM_HIP(hipMemcpy(c.cold_bm1, tbm + n, 8 * static_cast<size_t>(total), hipMemcpyDeviceToDevice));
      PM_HIP(hipMemcpy(c.cold_bm2, tbm + 2 * n, 8 * static_cast<size_t>(total), hipMemcpyDeviceToDevice));
    }
  }
  PM_HIP(hipDeviceSynchronize());
  (void)hipFree(tid);
  if (tbm) (void)hipFree(tbm);
  *next_slot = total;
  return PELEMP_OK;
}

// Move the live droplets of buffer cur, in sorted order, to the front of the
// other buffer (dead entries of the last pass are dropped); the cell
// offsets do not change and the permutation becomes the identity.
static int compact_storage(Ctx& c) {
  const int32_t* total = c.cstart + c.ncells;
  const int nxt = 1 - c.cur;
  const dim3 grid(1024), block(256);
  if (c.wsize == 8) {
    hipLaunchKernelGGL((gather_rows_kernel<double>), grid, block, 0, 0, total, c.perm,
                       static_cast<const double*>(c.fbuf[c.cur]), static_cast<double*>(c.fbuf[nxt]), c.nhot, c.cap);
  } else {
    hipLaunchKernelGGL((gather_rows_kernel<float>), grid, block, 0, 0, total, c.perm,
                       static_cast<const float*>(c.fbuf[c.cur]), static_cast<float*>(c.fbuf[nxt]), c.nhot, c.cap);
    hipLaunchKernelGGL((gather_rows_kernel<double>), grid, block, 0, 0, total, c.perm, c.mbuf[c.cur], c.mbuf[nxt],
                       c.nf, c.cap);
  }
  hipLaunchKernelGGL((gather_rows_kernel<int32_t>), grid, block, 0, 0, total, c.perm, c.slot[c.cur], c.slot[nxt], 1,
                     c.cap);
  // keys through the arrival scratch (the key array is not double-buffered)
  hipLaunchKernelGGL((gather_rows_kernel<uint32_t>), grid, block, 0, 0, total, c.perm, c.key,
                     reinterpret_cast<uint32_t*>(c.arrival), 1, c.cap);
  hipLaunchKernelGGL(identity_perm_kernel, grid, block, 0, 0, total, c.perm, c.n_entries);
  PM_HIP(hipGetLastError());
  int32_t n = 0;
  PM_HIP(hipMemcpy(&n, total, 4, hipMemcpyDeviceToHost));
  if (n > 0) PM_HIP(hipMemcpy(c.key, c.arrival, 4 * static_cast<size_t>(n), hipMemcpyDeviceToDevice));
  PM_HIP(hipDeviceSynchronize());
  c.cur = nxt;
  return PELEMP_OK;
}

// The FP32 and surrogate forms have no walls, so no wall films: a droplet
// handed in with a film height is refused.
static int check_no_film(const Ctx& c, const double* hr, int64_t m, int nsr) {
  if (c.form == PELEMP_FORM_FP64) return PELEMP_OK;
  for (int64_t i = 0; i < m; ++i)
    if (hr[nsr * i + 9 + c.nf] > 0.) {
      set_last_error("the FP32 and surrogate forms have no wall films (a droplet has a film height)");
      return PELEMP_E_BAD_CONFIG;
    }
  return PELEMP_OK;
}

// Append m droplets staged on the device in the AoS meaning (SI) after the
// storage entries of buffer cur, then sort.  new_next_id: the id counter
// after these droplets (at least the current one).
static int append_staged(Ctx& c, int64_t m, const double* dpos, const double* drd, const int64_t* did,
                         int64_t new_next_id) {
  if (m <= 0) return PELEMP_OK;
  PM_HIP(hipDeviceSynchronize());
  int32_t n_old = 0;
  PM_HIP(hipMemcpy(&n_old, c.n_entries, 4, hipMemcpyDeviceToHost));
  if (static_cast<int64_t>(n_old) + m > c.cap) {
    // entries deleted by the last pass still occupy storage: drop them
    int32_t total = 0;
    PM_HIP(hipMemcpy(&total, c.cstart + c.ncells, 4, hipMemcpyDeviceToHost));
    if (static_cast<int64_t>(total) + m > c.cap) {
      set_last_error("context capacity too small for the new droplets");
      return PELEMP_E_NOMEM;
    }
    const int st0 = compact_storage(c);
    if (st0) return st0;
    n_old = total;
  }
  int64_t next_id = 0, next_slot = 0;
  int st = read_counters(c, &next_id, &next_slot);
  if (st) return st;
  if (next_slot + m > c.slot_cap) {
    st = compact_slots(c, &next_slot);
    if (st) return st;
    if (next_slot + m > c.slot_cap) {
      set_last_error("context capacity too small for the new droplets");
      return PELEMP_E_NOMEM;
    }
  }
  const int nsr = 10 + c.nf;
  const double dxi[3] = {1. / c.geom.dx[0], 1. / c.geom.dx[1], 1. / c.geom.dx[2]};
  const dim3 grid(static_cast<unsigned>((m + 255) / 256)), block(256);
  if (c.form == PELEMP_FORM_FP64) {
    double* f = static_cast<double*>(c.fbuf[c.cur]) + n_old;
    hipLaunchKernelGGL((ingest_kernel<double, true>), grid, block, 0, 0, m, c.nf, nsr, dpos, drd, did, f, c.cap,
                       c.hot_brk ? 1 : 0, c.slot[c.cur] + n_old, next_slot, c.cold_id, c.cold_n0, c.cold_bm1,
                       c.cold_bm2, c.key + n_old, c.geom.plo[0], c.geom.plo[1], c.geom.plo[2], dxi[0], dxi[1],
                       dxi[2], c.ncell[0], c.ncell[1], c.ncell[2], c.ntile[0], c.ntile[1], c.ntile[2]);
  } else {
    double* dm = nullptr;
    PM_HIP(hipMalloc(&dm, 8 * m));
    RhoCoef rc{};
    for (int q = 0; q < 4 * PELEMP_MAX_FUEL; ++q) rc.a[q] = c.sd.rho_coef[q];
    hipLaunchKernelGGL(drop_mass_kernel, grid, block, 0, 0, m, c.nf, nsr, drd, rc, dm);
    float* f = static_cast<float*>(c.fbuf[c.cur]) + n_old;
    hipLaunchKernelGGL(ingest32_kernel, grid, block, 0, 0, m, c.nf, nsr, dpos, drd, dm, did, f,
                       c.mbuf[c.cur] + n_old, c.cap, c.slot[c.cur] + n_old, next_slot, c.cold_id, c.cold_n0,
                       c.cold_bm1, c.cold_bm2, c.key + n_old, c.geom.plo[0], c.geom.plo[1], c.geom.plo[2], dxi[0],
                       dxi[1], dxi[2], c.ncell[0], c.ncell[1], c.ncell[2], c.ntile[0], c.ntile[1], c.ntile[2]);
    PM_HIP(hipGetLastError());
    PM_HIP(hipDeviceSynchronize());
    (void)hipFree(dm);
  }
  PM_HIP(hipGetLastError());
  const int32_t n_new = n_old + static_cast<int32_t>(m);
  PM_HIP(hipMemcpy(c.n_entries, &n_new, 4, hipMemcpyHostToDevice));
  st = run_sort(c, 0, true);
  if (st) return st;
  PM_HIP(hipDeviceSynchronize());
  st = write_counters(c, std::max(next_id, new_next_id), next_slot + m);
  if (st) return st;
  int32_t total = 0;
  PM_HIP(hipMemcpy(&total, c.cstart + c.ncells, 4, hipMemcpyDeviceToHost));
  c.host_count = total;
  return PELEMP_OK;
}

// ---------------------------------------------------------------------------
// Spray jets (SprayJet.{H,cpp})

// interpolateROI (SprayJet.H:223-239); the bracketing index is kept inside
// the table (the reference reads before its start at time == time[0])
static double roi_interp(const pelemp_spray_jet& j, double time, const double* vals) {
  int i = 0;
  double ctime = j.roi_time[0];
  while (ctime < time && i < j.n_roi - 1) ctime = j.roi_time[++i];
  if (i < 1) i = 1;
  const double time1 = j.roi_time[i - 1];
  const double time2 = j.roi_time[i];
  const double invt = (time - time1) / (time2 - time1);
  const double val1 = vals[i - 1];
  const double val2 = vals[i];
  return val1 + (val2 - val1) * invt;
}

// jet_vel / mass_flow_rate (SprayJet.H:38-57), in the context's units
static double jet_velocity(const pelemp_spray_jet& j, double time) {
  return j.n_roi > 0 ? roi_interp(j, time, j.roi_vel) : j.jet_vel;
}
static double jet_mdot(const pelemp_spray_jet& j, double time) {
  return j.n_roi > 0 ? roi_interp(j, time, j.roi_mdot) : j.mass_flow;
}

// ChiSquared::init (Distributions.cpp:134-149)
static void chi_table(double* rvals) {
  const double xiend = 12.;
  const double dxi = xiend / 100.;
  const double rend = 1. - pm::exp(-xiend) * (1. + xiend * (1. + xiend * (0.5 + xiend / 6.)));
  for (int i = 0; i < kChiBins; i++) {
    const double xi = dxi * (i + 1);
    const double rval = 1. - pm::exp(-xi) * (1. + xi * (1. + xi * (0.5 + xi / 6.)));
    rvals[i] = rval / rend;
  }
}

// the distribution as sampled, and get_avg_dia (Distributions.cpp)
static InjDist make_dist(const pelemp_spray_jet& j, double* avg_dia) {
  InjDist d{};
  d.type = j.dist_type;
  d.p0 = j.dist_param[0];
  d.p1 = j.dist_param[1];
  *avg_dia = j.dist_param[0];
  if (j.dist_type == PELEMP_DIST_LOGNORMAL) {
    // LogNormal::init (Distributions.cpp:67-75)
    const double mean = j.dist_param[0], sd = j.dist_param[1];
    d.p0 = 2. * pm::log(mean) - 0.5 * pm::log(sd * sd + mean * mean);
    d.p1 = std::sqrt(std::max(-2. * pm::log(mean) + pm::log(sd * sd + mean * mean), 0.));
  } else if (j.dist_type == PELEMP_DIST_CHISQUARED) {
    chi_table(d.chi);
    // ChiSquared::get_avg_dia (Distributions.cpp:159-174)
    const double dmean = j.dist_param[0] / 3.;
    const double dxi = 12. / 100.;
    const double fact = 0.5;
    int curn = 0;
    double curr = d.chi[0];
    double curxi = 0.;
    while (fact > curr) {
      curn++;
      curr = d.chi[curn];
      curxi += dxi;
    }
    *avg_dia = curxi * dmean;
  }
  return d;
}

static int check_jet(const Ctx& c, pelemp_spray_jet& j, double* avg_dia) {
  if (j.dist_type < PELEMP_DIST_UNIFORM || j.dist_type > PELEMP_DIST_CHISQUARED) {
    set_last_error("unknown droplet distribution");
    return PELEMP_E_BAD_INJECTION;
  }
  if (!(j.spread_angle >= 0. && j.spread_angle <= kInjPI)) {
    set_last_error("Spread angle must be between 0 and 180");
    return PELEMP_E_BAD_INJECTION;
  }
  if (!(j.swirl_angle >= 0. && j.swirl_angle <= 0.5 * kInjPI)) {
    set_last_error("Swirl angle must be between 0 and 90");
    return PELEMP_E_BAD_INJECTION;
  }
  if (!(j.hollow_spread >= 0. && j.hollow_spread <= kInjPI)) {
    set_last_error("Hollow spread must be between 0 and 180");
    return PELEMP_E_BAD_INJECTION;
  }
  if (j.n_roi == 0 && (j.jet_vel < 0. || j.mass_flow < 0.)) {
    set_last_error("Jet u, T, and mdot must be greater than 0");
    return PELEMP_E_BAD_INJECTION;
  }
  if (j.n_roi < 0 || (j.n_roi > 0 && (j.n_roi < 2 || !j.roi_time || !j.roi_mdot || !j.roi_vel))) {
    set_last_error("rate-of-injection table needs at least two points");
    return PELEMP_E_ARG;
  }
  const double mag = std::sqrt(j.norm[0] * j.norm[0] + j.norm[1] * j.norm[1] + j.norm[2] * j.norm[2]);
  if (!(mag > 0.)) {
    set_last_error("jet normal is zero");
    return PELEMP_E_BAD_INJECTION;
  }
  for (int d = 0; d < 3; ++d) j.norm[d] /= mag;
  for (int d = 0; d < 3; ++d) {
    const double x = c.pelec ? j.cent[d] * cgs::kLen : j.cent[d];
    if (x > c.geom.phi[d] || x < c.geom.plo[d]) {
      set_last_error("Jet center outside of problem domain");
      return PELEMP_E_BAD_INJECTION;
    }
  }
  if (c.nf > 1) {
    double sumtest = 0.;
    for (int s = 0; s < c.nf; ++s) sumtest += j.jet_Y[s];
    if (std::fabs(1. - sumtest) > 1.E-8) {
      set_last_error("Jet mass fractions must sum to 1");
      return PELEMP_E_BAD_INJECTION;
    }
  }
  make_dist(j, avg_dia);
  return PELEMP_OK;
}

// Batches of candidates (inject_kernels.h) until the mass is reached or
// `limit` candidates are drawn.  Returns the droplets kept (also those past
// out_cap, which are counted but not staged), the injected mass, whether the
// limit stopped the loop and whether a droplet lies outside the domain.
struct InjOutcome {
  long long n_out = 0;
  double cur_mass = 0.;
  bool unfinished = false;
  bool outside = false;
};

template <int NF>
static int run_inject(const InjJet<NF>& jj, double inject_mass, long long limit, long long batch, long long out_cap,
                      long long next_id, double* dpos, double* drd, long long* did, double* dmass, int* dblk,
                      InjResult* dres, hipStream_t s, InjOutcome* out) {
  *out = InjOutcome{};
  if (!(0. < inject_mass)) return PELEMP_OK;  // the reference loop does not start
  InjResult r{};
  r.stop = -1;
  PM_HIP(hipMemcpyAsync(dres, &r, sizeof(r), hipMemcpyHostToDevice, s));
  long long k0 = 0, out_base = 0;
  while (true) {
    const long long nb = std::min(batch, limit - k0);
    const unsigned nblk = static_cast<unsigned>(nb / kInjBlock);
    hipLaunchKernelGGL((inj_mass_kernel<NF>), dim3(nblk), dim3(kInjBlock), 0, s, jj, k0, dmass, dblk);
    hipLaunchKernelGGL(inj_scan_kernel, dim3(1), dim3(64), 0, s, dmass, nb, dblk, static_cast<int>(nblk),
                       inject_mass, dres);
    hipLaunchKernelGGL((inj_write_kernel<NF>), dim3(nblk), dim3(kInjBlock), 0, s, jj, k0, dblk, out_base, out_cap,
                       next_id, dpos, drd, did, dres);
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpyAsync(&r, dres, sizeof(r), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    out->outside = out->outside || r.outside != 0;
    out->cur_mass = r.cur_mass;
    if (r.stop >= 0) {
      out->n_out = r.n_out;
      return PELEMP_OK;
    }
    out_base += r.batch_kept;
    k0 += nb;
    if (k0 >= limit) {
      out->n_out = out_base;
      out->unfinished = true;
      return PELEMP_OK;
    }
  }
}

// The droplet loop of sprayInjection on the device; droplets staged and
// appended.  jet: SI.
template <int NF>
static int inject_droplets(Ctx& c, const pelemp_spray_jet& j, double time, double dt, double inject_mass,
                           double num_ppp, double min_dia, double expected, double* cur_mass, int64_t* n_added,
                           hipStream_t s) {
  InjJet<NF> jj{};
  double avg_dia = 0.;
  jj.dist = make_dist(j, &avg_dia);
  for (int d = 0; d < 3; ++d) jj.cent[d] = j.cent[d];
  // transform_loc_vel's jet angles (SprayJet.H:175-181)
  const double norm_mag = std::sqrt(j.norm[0] * j.norm[0] + j.norm[1] * j.norm[1] + j.norm[2] * j.norm[2]);
  const double theta_jet = pm::acos(j.norm[2] / norm_mag);
  const double phi_jet = pm::atan2(j.norm[1] / norm_mag, j.norm[0] / norm_mag);
  jj.sp1 = pm::sin(phi_jet);
  jj.cp1 = pm::cos(phi_jet);
  jj.st1 = pm::sin(theta_jet);
  jj.ct1 = pm::cos(theta_jet);
  jj.spread = j.spread_angle;
  jj.swirl = j.swirl_angle;
  jj.hollow_spread = j.hollow_spread;
  jj.jet_dia = j.jet_dia;
  jj.hollow = j.hollow_spray;
  jj.umag = jet_velocity(j, time);
  jj.T = j.jet_T;
  // rho_part of an injected droplet (SprayInjection.H:129-139)
  double rho_part = 0.;
  if (NF > 1) {
    for (int s2 = 0; s2 < NF; ++s2) {
      jj.Y[s2] = j.jet_Y[s2];
      const double* a = &c.sd.rho_coef[4 * s2];
      rho_part += j.jet_Y[s2] / (a[0] + j.jet_T * (a[1] + j.jet_T * (a[2] + j.jet_T * a[3])));
    }
    rho_part = 1. / rho_part;
  } else {
    jj.Y[0] = 1.;
    const double* a = &c.sd.rho_coef[0];
    rho_part = a[0] + j.jet_T * (a[1] + j.jet_T * (a[2] + j.jet_T * a[3]));
  }
  if (!(rho_part > 0.)) {
    // parcel masses would not be positive: the reference loop never ends
    set_last_error("injection: non-positive liquid density at the jet temperature");
    return PELEMP_E_BAD_INJECTION;
  }
  jj.rho_part = rho_part;
  jj.min_dia = min_dia;
  jj.num_ppp = num_ppp;
  jj.dt = dt;
  jj.initial_bm2 = c.sd.do_breakup == 2 ? -1. : 0.;
  jj.seed = j.seed;
  jj.call = j.n_injections;
  for (int d = 0; d < 3; ++d) {
    jj.plo[d] = c.geom.plo[d];
    jj.dxi[d] = 1. / c.geom.dx[d];
    jj.ncell[d] = c.ncell[d];
  }
  // candidate limit: the reference loops until the mass is reached; a
  // distribution that keeps almost no candidates would never finish
  const double want = std::min(std::max(16. * expected + 65536., 1048576.), 6.7108864e7);
  const long long max_cand = static_cast<long long>(want);
  int64_t next_id = 0, next_slot = 0;
  int st = read_counters(c, &next_id, &next_slot);
  if (st) return st;
  long long out_cap = static_cast<long long>(std::min(2. * expected + 1024., static_cast<double>(c.cap) + 1.));
  const int nsr = 10 + NF;
  // candidates in batches of about the expected count (whole workgroups),
  // up to max_cand rounded up to whole workgroups
  const long long limit = (max_cand + kInjBlock - 1) / kInjBlock * kInjBlock;
  const double bw = std::min(std::max(1.05 * expected + 4096., 4096.), 1.6777216e7);
  const long long batch = std::min(limit, (static_cast<long long>(bw) + kInjBlock - 1) / kInjBlock * kInjBlock);
  double *dpos = nullptr, *drd = nullptr, *dmass = nullptr;
  long long* did = nullptr;
  int* dblk = nullptr;
  InjResult* dres = nullptr;
  auto release = [&]() {
    for (void* p : {static_cast<void*>(dpos), static_cast<void*>(drd), static_cast<void*>(did),
                    static_cast<void*>(dmass), static_cast<void*>(dblk), static_cast<void*>(dres)})
      if (p) (void)hipFree(p);
    dpos = drd = dmass = nullptr;
    did = nullptr;
    dblk = nullptr;
    dres = nullptr;
  };
  InjOutcome res{};
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (hipMalloc(&dpos, 8 * 3 * static_cast<size_t>(out_cap)) != hipSuccess ||
        hipMalloc(&drd, 8 * nsr * static_cast<size_t>(out_cap)) != hipSuccess ||
        hipMalloc(&did, 8 * static_cast<size_t>(out_cap)) != hipSuccess ||
        hipMalloc(&dmass, 8 * static_cast<size_t>(batch)) != hipSuccess ||
        hipMalloc(&dblk, 4 * static_cast<size_t>(batch / kInjBlock)) != hipSuccess ||
        hipMalloc(&dres, sizeof(InjResult)) != hipSuccess) {
      release();
      set_last_error("injection: device allocation failed");
      return PELEMP_E_NOMEM;
    }
    st = run_inject<NF>(jj, inject_mass, limit, batch, out_cap, next_id, dpos, drd, did, dmass, dblk, dres, s, &res);
    if (st) {
      release();
      return st;
    }
    if (res.n_out <= out_cap) break;
    // more droplets than staged: the same draws again with room for all
    release();
    out_cap = res.n_out;
  }
  if (res.unfinished) {
    release();
    set_last_error("injection: candidate limit reached before the injected mass");
    return PELEMP_E_BAD_INJECTION;
  }
  if (res.outside) {
    release();
    set_last_error("Bad injection particle");
    return PELEMP_E_BAD_INJECTION;
  }
  *cur_mass = res.cur_mass;
  *n_added = res.n_out;
  st = append_staged(c, res.n_out, dpos, drd, reinterpret_cast<const int64_t*>(did), next_id + res.n_out);
  release();
  return st;
}

// The domain check of updateParticles (SprayParticles.cpp:347-351 with
// check_bounds, SprayInterpolation.H:65-100) on droplets handed in: a
// droplet behind an outflow face has left the domain; one behind a
// reflective face is kept (the wall placement of film droplets leaves them a
// rounding error behind the wall); on a periodic axis the position must lie
// in [plo, phi].
int check_in_domain(const pelemp_geom& g, const double* pos, int64_t m) {
  for (int64_t i = 0; i < m; ++i)
    for (int d = 0; d < 3; ++d) {
      const double x = pos[3 * i + d];
      if (!(x == x)) {
        set_last_error("particle position is NaN");
        return PELEMP_E_LEFT_DOMAIN;
      }
      const bool below = x < g.plo[d], above = x > g.phi[d];
      if (!below && !above) continue;
      const int kind = below ? g.bndry_lo[d] : g.bndry_hi[d];
      if (kind == 1) continue;
      set_last_error(kind == 0 ? "particle outside the domain on a periodic axis (not wrapped)"
                               : "particle outside the domain (behind an outflow face)");
      return PELEMP_E_LEFT_DOMAIN;
    }
  return PELEMP_OK;
}

// real-component names of SprayParticleIO (SprayIO.cpp:9-33)
std::vector<std::string> spray_comp_names(const Ctx& c, int32_t nfuel_names, const char* const* fuel_names) {
  std::vector<std::string> names = {"xvel", "yvel", "zvel", "temperature", "diam"};
  for (int sp = 0; sp < c.nf; ++sp)
    names.push_back("spray_mf_" + ((fuel_names && sp < nfuel_names && fuel_names[sp]) ? std::string(fuel_names[sp])
                                                                                       : std::to_string(sp)));
  names.push_back("number_density");
  names.push_back("num_dens0");
  if (c.sd.do_breakup == 1) {
    names.push_back("Y_TAB");
    names.push_back("Ydot_TAB");
  } else if (c.sd.do_breakup == 2) {
    names.push_back("shed_mass");
    names.push_back("rt_time");
  } else {
    names.push_back("unused1");
    names.push_back("unused2");
  }
  names.push_back("wall_film_height");
  return names;
}

}  // namespace pelemp

using namespace pelemp;

extern "C" {

const char* pelemp_last_error(void) { return g_last_error.c_str(); }

int pelemp_spray_create(pelemp_spray_ctx** out, int32_t mech, int32_t form, const pelemp_spray_data* sd,
                        const pelemp_spray_comps* comps, const pelemp_geom* geom, int64_t max_particles,
                        int32_t apply_latent_shift) {
  if (!out || !sd || !comps || !geom || max_particles <= 0 || max_particles > 0x7ffffff0LL) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  *out = nullptr;
  const int ns = mech_ns(mech);
  if (ns < 0) {
    set_last_error("unknown mechanism");
    return PELEMP_E_ARG;
  }
  if (form != PELEMP_FORM_FP64 && form != PELEMP_FORM_FP32 && form != PELEMP_FORM_SURROGATE) {
    set_last_error("unsupported form");
    return PELEMP_E_ARG;
  }
  if (form == PELEMP_FORM_SURROGATE &&
      !((mech == PELEMP_MECH_DODECANE_AIR && sd->nfuel == 1) ||
        (mech == PELEMP_MECH_DECANE_DODECANE_AIR && sd->nfuel == 2))) {
    set_last_error("the surrogate form supports mechanisms 0 (one fuel) and 1 (two fuels)");
    return PELEMP_E_BAD_CONFIG;
  }
  if (sd->nfuel < 1 || sd->nfuel > PELEMP_MAX_FUEL) {
    set_last_error("nfuel out of range");
    return PELEMP_E_BAD_CONFIG;
  }
  for (int i = 0; i < sd->nfuel; ++i) {
    if (sd->indx[i] < 0 || sd->indx[i] >= ns || sd->dep_indx[i] < 0 || sd->dep_indx[i] >= ns) {
      set_last_error("fuel species index not in the mechanism");  // SpraySetup.cpp:272-277
      return PELEMP_E_BAD_CONFIG;
    }
    // the kernels address the fuel species at compile time: every compiled
    // mechanism lists its fuels last, and the liquid's species are those
    // fuels in order.  The vapour may go to any species (dep_fuel_species,
    // SpraySetup.cpp:103-125), except in the surrogate form, whose network
    // sees the fuel's own gas-phase mass fraction.
    if (sd->indx[i] != ns - sd->nfuel + i ||
        (form == PELEMP_FORM_SURROGATE && sd->dep_indx[i] != sd->indx[i])) {
      set_last_error("fuel species must be the last nfuel species of the mechanism, in order "
                     "(and, in the surrogate form, deposit into themselves)");
      return PELEMP_E_BAD_CONFIG;
    }
  }
  if (sd->do_breakup < 0 || sd->do_breakup > 2) {
    set_last_error("do_breakup must be 0 (none), 1 (TAB) or 2 (KH-RT)");  // SpraySetup.cpp:138-155
    return PELEMP_E_BAD_CONFIG;
  }
  if ((sd->do_breakup != 0 || sd->do_splash != 0) && form != PELEMP_FORM_FP64) {
    set_last_error("breakup / splash / wall films are supported by the FP64 form only");
    return PELEMP_E_BAD_CONFIG;
  }
  if (sd->do_splash != 0 && !(sd->sigma > 0. && sd->wall_T > 0. && sd->theta_c > 0.)) {
    set_last_error("splash needs fuel_sigma, wall_temp and contact_angle");  // SpraySetup.cpp:157-183
    return PELEMP_E_BAD_CONFIG;
  }
  if (sd->do_breakup != 0 && !(sd->sigma > 0.)) {
    set_last_error("breakup needs fuel_sigma");  // SpraySetup.cpp:157-165
    return PELEMP_E_BAD_CONFIG;
  }
  {
    // the deposit writes each source quantity to its own component
    // (SprayComps: momentum, density, fuel species, energy)
    std::vector<int> sc = {comps->momSrcIndx, comps->momSrcIndx + 1, comps->momSrcIndx + 2, comps->rhoSrcIndx,
                           comps->engSrcIndx};
    for (int i = 0; i < sd->nfuel; ++i) sc.push_back(comps->specSrcIndx + i);
    std::sort(sc.begin(), sc.end());
    if (sc.front() < 0 || std::adjacent_find(sc.begin(), sc.end()) != sc.end()) {
      set_last_error("SprayComps: the source components overlap or are negative");
      return PELEMP_E_BAD_CONFIG;
    }
  }
  auto* h = new pelemp_spray_ctx();
  Ctx& c = h->c;
  c.mech = mech;
  c.form = form;
  c.nf = sd->nfuel;
  c.ns = ns;
  c.sd = *sd;
  c.comps = *comps;
  c.geom = *geom;
  mech_mw(mech, c.mw);
  if (apply_latent_shift) {
    // latent = ref_latent - h_gas,fuel(ref_T)   (SpraySetup.cpp:284-290)
    std::vector<double> hh(ns);
    mech_h(mech, sd->ref_T, hh.data());
    for (int i = 0; i < c.nf; ++i) c.sd.latent[i] = sd->ref_latent[i] - hh[sd->indx[i]];
  }
  for (int d = 0; d < 3; ++d) {
    c.ncell[d] = geom->domain_hi[d] - geom->domain_lo[d] + 1;
    if (c.ncell[d] <= 0 || c.ncell[d] % kTile != 0) {
      delete h;
      set_last_error("domain extent must be a positive multiple of 4 in every direction");
      return PELEMP_E_ARG;
    }
    c.ntile[d] = c.ncell[d] / kTile;
  }
  c.ncells = c.ncell[0] * c.ncell[1] * c.ncell[2];
  c.ntiles = c.ntile[0] * c.ntile[1] * c.ntile[2];
  c.cap = max_particles;
  // the branch kernel (films, TAB, wall reflection / splash) runs when
  // breakup or splash is on or a face is a reflective wall; the breakup
  // state is then carried in the hot fields
  bool reflective = false;
  for (int d = 0; d < 3; ++d) reflective = reflective || geom->bndry_lo[d] == 1 || geom->bndry_hi[d] == 1;
  c.branches = (form == PELEMP_FORM_FP64) && (sd->do_breakup != 0 || sd->do_splash != 0 || reflective);
  c.hot_brk = c.branches;
  c.slot_cap = c.branches ? 4 * c.cap : c.cap;
  c.params.max_parcel_size = 100.;
  c.params.breakup_parcel_factor = 0.5;
  c.params.khrt_B0 = 0.61;
  c.params.khrt_B1 = 7.;
  c.params.khrt_C3 = 1.;
  c.params.rng_seed = 0;
  c.params.avg_inject_mass = 0.;
  if (form == PELEMP_FORM_FP64) {
    c.nhot = c.hot_brk ? (13 + c.nf) : (10 + c.nf);
    c.wsize = 8;
  } else {
    c.nhot = Fields32::NHOT;
    c.wsize = 4;
  }
  int st = PELEMP_OK;
  const size_t fb = c.wsize * c.nhot * static_cast<size_t>(c.cap);
  st |= alloc(c, &c.fbuf[0], fb);
  st |= alloc(c, &c.fbuf[1], fb);
  if (form != PELEMP_FORM_FP64) {
    st |= alloc(c, reinterpret_cast<void**>(&c.mbuf[0]), 8 * static_cast<size_t>(c.nf) * c.cap);
    st |= alloc(c, reinterpret_cast<void**>(&c.mbuf[1]), 8 * static_cast<size_t>(c.nf) * c.cap);
  }
  st |= alloc(c, reinterpret_cast<void**>(&c.slot[0]), 4 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.slot[1]), 4 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.perm), 4 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.cstart), 4 * (static_cast<size_t>(c.ncells) + 1));
  st |= alloc(c, reinterpret_cast<void**>(&c.cnt), 4 * static_cast<size_t>(c.ncells));
  st |= alloc(c, reinterpret_cast<void**>(&c.arrival), 4 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.long_list), sizeof(int2) * (c.cap / (kWaveSortMax + 1) + 1));
  // mov_lo, stay_cnt, mov_hi in one allocation so one memset clears them
  st |= alloc(c, reinterpret_cast<void**>(&c.mov_lo), 3 * 4 * static_cast<size_t>(c.ncells));
  st |= alloc(c, reinterpret_cast<void**>(&c.key), 4 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.block_sum),
              4 * (static_cast<size_t>(c.ncells) / kScanBlock + 2));
  st |= alloc(c, reinterpret_cast<void**>(&c.scalars), 64);
  st |= alloc(c, reinterpret_cast<void**>(&c.dscratch), 64);
  st |= alloc(c, reinterpret_cast<void**>(&c.cold_id), 8 * c.slot_cap);
  if (c.branches) st |= alloc_branch_buffers(c);
  st |= alloc(c, reinterpret_cast<void**>(&c.cold_n0), 8 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.cold_bm1), 8 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.cold_bm2), 8 * c.cap);
  st |= alloc(c, reinterpret_cast<void**>(&c.count_log), 4 * Ctx::kCountLog);
  // heavy tiles: a tile of n > kSegDroplets droplets has ceil(n / kSegDroplets)
  // < 2 n / kSegDroplets segments, so the droplet capacity bounds the lists
  c.split_items_cap = static_cast<int32_t>(c.cap / kSegDroplets + 1);
  c.split_list_cap = c.split_items_cap;
  c.split_slots_cap = 2 * (c.cap / kSegDroplets) + 2;
  st |= alloc(c, reinterpret_cast<void**>(&c.sp.tile_slot), 4 * static_cast<size_t>(c.ntiles));
  st |= alloc(c, reinterpret_cast<void**>(&c.sp.items), sizeof(int2) * static_cast<size_t>(c.split_items_cap));
  st |= alloc(c, reinterpret_cast<void**>(&c.sp.list), sizeof(int4) * static_cast<size_t>(c.split_list_cap));
  st |= alloc(c, reinterpret_cast<void**>(&c.sp.part),
              8 * static_cast<size_t>(c.split_slots_cap) * kTileCells * (5 + c.nf));
  st |= alloc(c, reinterpret_cast<void**>(&c.sp.seg_stay), 4 * static_cast<size_t>(c.split_slots_cap) * kTileCells);
  if (form == PELEMP_FORM_SURROGATE) {
    st |= alloc(c, &c.surr_w, sizeof(SurrWeights));
    st |= alloc(c, reinterpret_cast<void**>(&c.surr_cnt), 2 * sizeof(unsigned long long));
  }
  if (st != PELEMP_OK) {
    free_ctx(c);
    delete h;
    return PELEMP_E_NOMEM;
  }
  c.stay_cnt = c.mov_lo + c.ncells;
  c.mov_hi = c.mov_lo + 2 * static_cast<size_t>(c.ncells);
  c.n_entries = c.scalars;
  c.total = c.scalars + 1;
  c.err = c.scalars + 2;
  c.n_long = c.scalars + 3;
  c.sp.counters = c.scalars + 4;
  if (c.surr_cnt && hipMemset(c.surr_cnt, 0, 2 * sizeof(unsigned long long)) != hipSuccess) st = PELEMP_E_HIP;
  if (c.sb_ids) {
    // counters of new droplets: next id, next slot, children of the last pass
    const int64_t ids[4] = {1, 0, 0, 0};
    if (hipMemcpy(c.sb_ids, ids, sizeof(ids), hipMemcpyHostToDevice) != hipSuccess) st = PELEMP_E_HIP;
  }
  if (st != PELEMP_OK || hipMemset(c.scalars, 0, 64) != hipSuccess ||
      hipMemset(c.sp.tile_slot, 0xff, 4 * static_cast<size_t>(c.ntiles)) != hipSuccess ||
      hipMemset(c.cstart, 0, 4 * (static_cast<size_t>(c.ncells) + 1)) != hipSuccess) {
    free_ctx(c);
    delete h;
    set_last_error("hipMemset failed");
    return PELEMP_E_HIP;
  }
  *out = h;
  return PELEMP_OK;
}

int pelemp_spray_create_pelec(pelemp_spray_ctx** out, int32_t mech, int32_t form, const pelemp_spray_data* sd_cgs,
                              const pelemp_spray_comps* comps, const pelemp_geom* geom_cgs, int64_t max_particles) {
  if (!out || !sd_cgs || !geom_cgs) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  if (form != PELEMP_FORM_FP64) {
    set_last_error("the PeleC flavour is supported by the FP64 form only");
    return PELEMP_E_BAD_CONFIG;
  }
  pelemp_spray_data sd = *sd_cgs;
  pelemp_geom g = *geom_cgs;
  cgs::spray_data_to_si(sd);
  cgs::geom_to_si(g);
  // the deck's reference latent heat is shifted in SI (SpraySetup.cpp:284-290)
  const int st = pelemp_spray_create(out, mech, form, &sd, comps, &g, max_particles, 1);
  if (st != PELEMP_OK) return st;
  Ctx& c = (*out)->c;
  c.pelec = true;
  cgs::source_scale(c.nf, c.src_scale);
  return PELEMP_OK;
}

int pelemp_spray_destroy(pelemp_spray_ctx* h) {
  if (!h) return PELEMP_OK;
  (void)hipDeviceSynchronize();
  free_ctx(h->c);
  delete h;
  return PELEMP_OK;
}

int pelemp_spray_get_data(const pelemp_spray_ctx* h, pelemp_spray_data* out) {
  if (!h || !out) return PELEMP_E_ARG;
  *out = h->c.sd;
  return PELEMP_OK;
}

int pelemp_spray_set_particles(pelemp_spray_ctx* h, int64_t n, const double* pos, const double* rdata,
                               const int64_t* id) {
  if (!h || n < 0 || (n > 0 && (!pos || !rdata || !id))) return PELEMP_E_ARG;
  Ctx& c = h->c;
  const int nsr = 10 + c.nf;
  // keep only active particles (id > 0); when all are active and no unit
  // conversion is needed, the caller's arrays are used as they are
  bool all_active = true;
  for (int64_t i = 0; i < n && all_active; ++i) all_active = id[i] > 0;
  std::vector<double> hpv, hrv;
  std::vector<int64_t> hidv;
  const double* hp = pos;
  const double* hr = rdata;
  const int64_t* hid = id;
  int64_t m = n;
  if (!all_active || c.pelec) {
    hpv.reserve(3 * n);
    hrv.reserve(nsr * n);
    hidv.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
      if (id[i] <= 0) continue;
      hpv.insert(hpv.end(), pos + 3 * i, pos + 3 * i + 3);
      hrv.insert(hrv.end(), rdata + nsr * i, rdata + nsr * i + nsr);
      hidv.push_back(id[i]);
      if (c.pelec) cgs::droplet_to_si(&hpv[hpv.size() - 3], &hrv[hrv.size() - nsr], c.nf, c.sd.do_breakup);
    }
    hp = hpv.data();
    hr = hrv.data();
    hid = hidv.data();
    m = static_cast<int64_t>(hidv.size());
  }
  if (m > c.cap) {
    set_last_error("more particles than the context capacity");
    return PELEMP_E_ARG;
  }
  if (const int st = check_in_domain(c.geom, hp, m)) return st;
  if (const int st = check_no_film(c, hr, m, nsr)) return st;
  double *dpos = nullptr, *drd = nullptr;
  int64_t* did = nullptr;
  PM_HIP(hipMalloc(&dpos, 8 * std::max<int64_t>(1, 3 * m)));
  PM_HIP(hipMalloc(&drd, 8 * std::max<int64_t>(1, nsr * m)));
  PM_HIP(hipMalloc(&did, 8 * std::max<int64_t>(1, m)));
  if (m > 0) {
    PM_HIP(hipMemcpy(dpos, hp, 8 * 3 * m, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(drd, hr, 8 * nsr * m, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(did, hid, 8 * m, hipMemcpyHostToDevice));
  }
  c.cur = 0;
  c.nslots = 0;
  c.npass = 0;
  const double dxi[3] = {1. / c.geom.dx[0], 1. / c.geom.dx[1], 1. / c.geom.dx[2]};
  if (m > 0) {
    const dim3 grid(static_cast<unsigned>((m + 255) / 256)), block(256);
    if (c.form == PELEMP_FORM_FP64) {
      hipLaunchKernelGGL((ingest_kernel<double, true>), grid, block, 0, 0, m, c.nf, nsr, dpos, drd, did,
                         static_cast<double*>(c.fbuf[0]), c.cap, c.hot_brk ? 1 : 0, c.slot[0], c.nslots,
                         c.cold_id, c.cold_n0, c.cold_bm1, c.cold_bm2, c.key, c.geom.plo[0], c.geom.plo[1],
                         c.geom.plo[2], dxi[0], dxi[1], dxi[2], c.ncell[0], c.ncell[1], c.ncell[2],
                         c.ntile[0], c.ntile[1], c.ntile[2]);
    } else {
      // droplet mass pi/6 rho_l(T) d^3 with the mixture liquid density
      // 1/sum(Y/rhoL(T)) (Drag.H:190-199 at T below the boiling point)
      std::vector<double> hm(m);
      for (int64_t i = 0; i < m; ++i) {
        const double* r = &hr[nsr * i];
        const double T = r[3];
        double inv = 0.;
        for (int s = 0; s < c.nf; ++s) {
          const double* a = &c.sd.rho_coef[4 * s];
          inv += r[5 + s] / (a[0] + T * (a[1] + T * (a[2] + T * a[3])));
        }
        hm[i] = kPIc / 6. * (1. / inv) * (r[4] * r[4] * r[4]);
      }
      double* dm = nullptr;
      PM_HIP(hipMalloc(&dm, 8 * m));
      PM_HIP(hipMemcpy(dm, hm.data(), 8 * m, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(ingest32_kernel, grid, block, 0, 0, m, c.nf, nsr, dpos, drd, dm, did,
                         static_cast<float*>(c.fbuf[0]), c.mbuf[0], c.cap, c.slot[0], c.nslots, c.cold_id,
                         c.cold_n0, c.cold_bm1, c.cold_bm2, c.key, c.geom.plo[0], c.geom.plo[1], c.geom.plo[2],
                         dxi[0], dxi[1], dxi[2], c.ncell[0], c.ncell[1], c.ncell[2], c.ntile[0], c.ntile[1],
                         c.ntile[2]);
      PM_HIP(hipDeviceSynchronize());
      (void)hipFree(dm);
    }
    PM_HIP(hipGetLastError());
  }
  c.nslots = m;
  int64_t max_id = 0;
  for (int64_t i = 0; i < m; ++i) max_id = std::max(max_id, hid[i]);
  c.next_id = max_id + 1;
  if (c.branches) {
    // counters of new droplets: next id (past the largest id handed in),
    // next slot, children of the last pass
    const int64_t ids[4] = {max_id + 1, m, 0, 0};
    PM_HIP(hipMemcpy(c.sb_ids, ids, sizeof(ids), hipMemcpyHostToDevice));
  }
  // the sort reads the entry count from cstart[ncells]
  const int32_t m32 = static_cast<int32_t>(m);
  PM_HIP(hipMemcpy(c.cstart + c.ncells, &m32, 4, hipMemcpyHostToDevice));
  PM_HIP(hipMemset(c.err, 0, 4));
  const int st = run_sort(c, 0);
  PM_HIP(hipDeviceSynchronize());
  (void)hipFree(dpos);
  (void)hipFree(drd);
  (void)hipFree(did);
  c.host_count = m;
  return st;
}

int pelemp_spray_num_particles(pelemp_spray_ctx* h, int64_t* n) {
  if (!h || !n) return PELEMP_E_ARG;
  int32_t t = 0;
  PM_HIP(hipDeviceSynchronize());
  PM_HIP(hipMemcpy(&t, h->c.cstart + h->c.ncells, 4, hipMemcpyDeviceToHost));
  *n = t;
  h->c.host_count = t;
  return PELEMP_OK;
}

int pelemp_spray_get_particles(pelemp_spray_ctx* h, int64_t* n, double* pos, double* rdata, int64_t* id) {
  if (!h || !n) return PELEMP_E_ARG;
  Ctx& c = h->c;
  int64_t m = 0;
  int st = pelemp_spray_num_particles(h, &m);
  if (st) return st;
  if (m > *n) {
    *n = m;
    set_last_error("output capacity too small");
    return PELEMP_E_ARG;
  }
  const int nsr = 10 + c.nf;
  double *dpos = nullptr, *drd = nullptr;
  int64_t* did = nullptr;
  PM_HIP(hipMalloc(&dpos, 8 * std::max<int64_t>(1, 3 * m)));
  PM_HIP(hipMalloc(&drd, 8 * std::max<int64_t>(1, nsr * m)));
  PM_HIP(hipMalloc(&did, 8 * std::max<int64_t>(1, m)));
  if (m > 0) {
    const dim3 grid(static_cast<unsigned>((m + 255) / 256)), block(256);
    if (c.form == PELEMP_FORM_FP64) {
      hipLaunchKernelGGL((gather_kernel<double, true>), grid, block, 0, 0, c.cstart + c.ncells, c.nf, nsr, c.perm,
                         static_cast<const double*>(c.fbuf[c.cur]), c.cap, c.hot_brk ? 1 : 0, c.slot[c.cur],
                         c.cold_id, c.cold_n0, c.cold_bm1, c.cold_bm2, c.key, dpos, drd, did, c.geom.plo[0],
                         c.geom.plo[1], c.geom.plo[2], c.geom.dx[0], c.geom.dx[1], c.geom.dx[2], c.ntile[0],
                         c.ntile[1], c.ntile[2]);
    } else {
      hipLaunchKernelGGL(gather32_kernel, grid, block, 0, 0, c.cstart + c.ncells, c.nf, nsr, c.perm,
                         static_cast<const float*>(c.fbuf[c.cur]), c.mbuf[c.cur], c.cap, c.slot[c.cur],
                         c.cold_id, c.cold_n0, c.cold_bm1, c.cold_bm2, c.key, dpos, drd, did, c.geom.plo[0],
                         c.geom.plo[1], c.geom.plo[2], c.geom.dx[0], c.geom.dx[1], c.geom.dx[2], c.ntile[0],
                         c.ntile[1], c.ntile[2]);
    }
    PM_HIP(hipGetLastError());
    PM_HIP(hipDeviceSynchronize());
    if (pos) PM_HIP(hipMemcpy(pos, dpos, 8 * 3 * m, hipMemcpyDeviceToHost));
    if (rdata) PM_HIP(hipMemcpy(rdata, drd, 8 * nsr * m, hipMemcpyDeviceToHost));
    if (id) PM_HIP(hipMemcpy(id, did, 8 * m, hipMemcpyDeviceToHost));
    if (c.pelec && pos && rdata) {
      for (int64_t i = 0; i < m; ++i) cgs::droplet_to_cgs(pos + 3 * i, rdata + nsr * i, c.nf, c.sd.do_breakup);
    }
  }
  (void)hipFree(dpos);
  (void)hipFree(drd);
  (void)hipFree(did);
  *n = m;
  return PELEMP_OK;
}

int pelemp_spray_add_particles(pelemp_spray_ctx* h, int64_t n, const double* pos, const double* rdata,
                               const int64_t* id) {
  if (!h || n < 0 || (n > 0 && (!pos || !rdata || !id))) return PELEMP_E_ARG;
  Ctx& c = h->c;
  const int nsr = 10 + c.nf;
  std::vector<double> hp, hr;
  std::vector<int64_t> hid;
  int64_t max_id = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (id[i] <= 0) continue;
    hp.insert(hp.end(), pos + 3 * i, pos + 3 * i + 3);
    hr.insert(hr.end(), rdata + nsr * i, rdata + nsr * i + nsr);
    hid.push_back(id[i]);
    max_id = std::max(max_id, id[i]);
    if (c.pelec) cgs::droplet_to_si(&hp[hp.size() - 3], &hr[hr.size() - nsr], c.nf, c.sd.do_breakup);
  }
  const int64_t m = static_cast<int64_t>(hid.size());
  if (m == 0) return PELEMP_OK;
  if (const int st = check_in_domain(c.geom, hp.data(), m)) return st;
  if (const int st = check_no_film(c, hr.data(), m, nsr)) return st;
  double *dpos = nullptr, *drd = nullptr;
  int64_t* did = nullptr;
  PM_HIP(hipMalloc(&dpos, 8 * 3 * m));
  PM_HIP(hipMalloc(&drd, 8 * nsr * m));
  PM_HIP(hipMalloc(&did, 8 * m));
  PM_HIP(hipMemcpy(dpos, hp.data(), 8 * 3 * m, hipMemcpyHostToDevice));
  PM_HIP(hipMemcpy(drd, hr.data(), 8 * nsr * m, hipMemcpyHostToDevice));
  PM_HIP(hipMemcpy(did, hid.data(), 8 * m, hipMemcpyHostToDevice));
  const int st = append_staged(c, m, dpos, drd, did, max_id + 1);
  (void)hipFree(dpos);
  (void)hipFree(drd);
  (void)hipFree(did);
  return st;
}

int pelemp_spray_jet_check(const pelemp_spray_ctx* h, pelemp_spray_jet* jet, double* avg_dia) {
  if (!h || !jet) return PELEMP_E_ARG;
  double a = 0.;
  const int st = check_jet(h->c, *jet, &a);
  if (avg_dia) *avg_dia = a;
  return st;
}

int pelemp_spray_read_roi(const char* path, double rho_part, double cd, double jet_dia, double mass_conv,
                          int32_t* n, double* time, double* mdot, double* vel, double* start_time,
                          double* end_time, double* max_jet_vel) {
  if (!path || !n || *n < 0) return PELEMP_E_ARG;
  std::ifstream infile(path);
  if (!infile.good()) {
    set_last_error("ROI file does not exists");
    return PELEMP_E_ARG;
  }
  std::stringstream buf;
  buf << infile.rdbuf();
  std::istringstream iss(buf.str());
  std::string line;
  std::getline(iss, line);  // header
  const double jet_area = 0.25 * kInjPI * pm::pow(jet_dia, 2.);
  const double tconv = 1.E-3;  // ms -> s
  double maxv = 0., t_start = 0., t_end = 1.E8, prev_time = 0.;
  bool started = false;
  std::vector<double> tt, mm, vv;
  while (std::getline(iss, line)) {
    for (char& ch : line)
      if (ch == ';') ch = ' ';
    std::istringstream sin(line);
    double t = 0., md = 0.;
    sin >> t;
    sin >> md;
    t *= tconv;
    md *= mass_conv;
    if (md > 0.) {
      if (!started) {
        started = true;
        t_start = t;
        tt.push_back(prev_time);
        mm.push_back(0.);
        vv.push_back(0.);
      }
      const double v = md / (rho_part * jet_area * cd);
      maxv = std::max(maxv, v);
      tt.push_back(t);
      mm.push_back(md);
      vv.push_back(v);
      t_end = t;
    } else if (started) {
      t_end = t;
      tt.push_back(t);
      mm.push_back(0.);
      vv.push_back(0.);
      break;
    }
    prev_time = t;
  }
  const int32_t np = static_cast<int32_t>(tt.size());
  if (np > *n) {
    *n = np;
    set_last_error("ROI table capacity too small");
    return PELEMP_E_ARG;
  }
  for (int32_t i = 0; i < np; ++i) {
    if (time) time[i] = tt[i];
    if (mdot) mdot[i] = mm[i];
    if (vel) vel[i] = vv[i];
  }
  *n = np;
  if (start_time) *start_time = t_start;
  if (end_time) *end_time = t_end;
  if (max_jet_vel) *max_jet_vel = maxv;
  return PELEMP_OK;
}

int pelemp_spray_inject(pelemp_spray_ctx* h, pelemp_spray_jet* jet, double time, double sim_dt, int64_t* n_injected,
                        void* stream) {
  if (!h || !jet) return PELEMP_E_ARG;
  Ctx& c = h->c;
  if (n_injected) *n_injected = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // jet_active (SprayJet.H:71-86); the checks on a copy, so the caller's
  // jet (normalised once by pelemp_spray_jet_check) is not changed
  double avg_dia = 0.;
  {
    pelemp_spray_jet chk = *jet;
    const int st0 = check_jet(c, chk, &avg_dia);
    if (st0) return st0;
  }
  int st = PELEMP_OK;
  if (jet->jet_dia < 0.) {
    set_last_error("Jet diameter is not set");
    return PELEMP_E_BAD_INJECTION;
  }
  if (avg_dia < 0.) {
    set_last_error("Average droplet diameter is not set");
    return PELEMP_E_BAD_INJECTION;
  }
  if (jet->jet_T < 0.) {
    set_last_error("Droplet temperature is not set");
    return PELEMP_E_BAD_INJECTION;
  }
  const bool active = !(time < jet->start_time || time > jet->end_time);
  // SprayInjection.H:17-22
  if (!active || sim_dt <= 0. || jet_velocity(*jet, time) < 0.) return PELEMP_OK;
  // the jet in SI (a PeleC jet is CGS); rates evaluated in the jet's units
  pelemp_spray_jet j = *jet;
  double mdot = jet_mdot(*jet, time);
  double umag = jet_velocity(*jet, time);
  if (c.pelec) {
    cgs::jet_to_si(j);
    mdot *= cgs::kMass;
    umag *= cgs::kLen;
    make_dist(j, &avg_dia);  // the SI mean diameter
  }
  j.n_roi = 0;  // rates below are the evaluated ones
  j.jet_vel = umag;
  // SprayInjection.H:28-66
  double rho_avg = 0.;
  for (int spf = 0; spf < c.nf; ++spf) {
    const double* a = &c.sd.rho_coef[4 * spf];
    const double Y = c.nf > 1 ? j.jet_Y[spf] : 1.;
    rho_avg += Y / (a[0] + j.jet_T * (a[1] + j.jet_T * (a[2] + j.jet_T * a[3])));
  }
  rho_avg = 1. / rho_avg;
  const double Pi_six = kInjPI / 6.;
  const double avg_mass = Pi_six * rho_avg * pm::pow(avg_dia, 3.);
  const double min_dia = pm::cbrt(6. * kMinMass / (kInjPI * rho_avg));
  if (avg_dia < min_dia || avg_mass < kMinMass) {
    set_last_error("Average droplet size too small, floating point issues expected");
    return PELEMP_E_BAD_INJECTION;
  }
  double dt = sim_dt;
  double inject_mass = mdot * sim_dt;
  if (j.sum_inj_mass > 0.) {
    dt += j.sum_inj_time;
    inject_mass += j.sum_inj_mass;
  }
  double num_ppp = j.num_ppp;
  if (num_ppp < 0.) {
    num_ppp = inject_mass / avg_mass;
    j.num_ppp = num_ppp;
  }
  if (inject_mass / (num_ppp * avg_mass) < j.min_parcel) {
    j.sum_inj_mass = inject_mass;
    j.sum_inj_time = dt;
  } else {
    const double expected = inject_mass / (num_ppp * avg_mass);
    double cur_mass = 0.;
    int64_t added = 0;
    switch (c.nf) {
      case 1: st = inject_droplets<1>(c, j, time, dt, inject_mass, num_ppp, min_dia, expected, &cur_mass, &added, s); break;
      case 2: st = inject_droplets<2>(c, j, time, dt, inject_mass, num_ppp, min_dia, expected, &cur_mass, &added, s); break;
      case 6: st = inject_droplets<6>(c, j, time, dt, inject_mass, num_ppp, min_dia, expected, &cur_mass, &added, s); break;
      default: set_last_error("unsupported fuel count"); return PELEMP_E_BAD_CONFIG;
    }
    if (st) return st;
    j.n_injections += 1;
    // SprayInjection.H:160-180
    const double est_mdot = cur_mass / dt;
    if (est_mdot - mdot > 0.5 * mdot) j.min_parcel += 1.;
    j.total_inj_mass += cur_mass;
    j.total_inj_time += dt;
    j.sum_inj_mass = -1.;
    j.sum_inj_time = -1.;
    if (n_injected) *n_injected = added;
  }
  // the state back into the caller's jet
  if (c.pelec) {
    cgs::jet_state_to_cgs(j, *jet);
  } else {
    jet->num_ppp = j.num_ppp;
    jet->sum_inj_mass = j.sum_inj_mass;
    jet->sum_inj_time = j.sum_inj_time;
    jet->min_parcel = j.min_parcel;
    jet->total_inj_mass = j.total_inj_mass;
    jet->total_inj_time = j.total_inj_time;
    jet->n_injections = j.n_injections;
  }
  return PELEMP_OK;
}

int pelemp_spray_uniform_init(pelemp_spray_ctx* h, const int32_t num_part[3], const double vel[3], double dia,
                              double T, const double* Y, double num_ppp) {
  if (!h || !num_part || !vel || !Y) return PELEMP_E_ARG;
  Ctx& c = h->c;
  if (num_part[0] <= 0 || num_part[1] <= 0 || num_part[2] <= 0) return PELEMP_E_ARG;
  const long long n = static_cast<long long>(num_part[0]) * num_part[1] * num_part[2];
  int64_t n_old = 0;
  int st = pelemp_spray_num_particles(h, &n_old);
  if (st) return st;
  if (n > c.cap) {
    set_last_error("context capacity too small for the new droplets");
    return PELEMP_E_NOMEM;
  }
  const double f = c.pelec ? cgs::kLen : 1.;
  double dxp[3];
  for (int d = 0; d < 3; ++d) dxp[d] = (c.geom.phi[d] - c.geom.plo[d]) / static_cast<double>(num_part[d]);
  const double v[3] = {vel[0] * f, vel[1] * f, vel[2] * f};
  const double dia_si = dia * f;
  const int nsr = 10 + c.nf;
  int64_t next_id = 0, next_slot = 0;
  st = read_counters(c, &next_id, &next_slot);
  if (st) return st;
  double *dpos = nullptr, *drd = nullptr, *dY = nullptr;
  long long* did = nullptr;
  int* dout = nullptr;
  PM_HIP(hipMalloc(&dpos, 8 * 3 * static_cast<size_t>(n)));
  PM_HIP(hipMalloc(&drd, 8 * nsr * static_cast<size_t>(n)));
  PM_HIP(hipMalloc(&did, 8 * static_cast<size_t>(n)));
  PM_HIP(hipMalloc(&dY, 8 * PELEMP_MAX_FUEL));
  PM_HIP(hipMalloc(&dout, 4));
  PM_HIP(hipMemcpy(dY, Y, 8 * c.nf, hipMemcpyHostToDevice));
  PM_HIP(hipMemset(dout, 0, 4));
  const double ib2 = c.sd.do_breakup == 2 ? -1. : 0.;
  const dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
  auto run = [&](auto nft) {
    constexpr int NF = decltype(nft)::value;
    hipLaunchKernelGGL((lattice_kernel<NF>), grid, block, 0, 0, n, num_part[0], num_part[1], dxp[0], dxp[1], dxp[2],
                       v[0], v[1], v[2], dia_si, T, dY, num_ppp, ib2, next_id, c.geom.plo[0], c.geom.plo[1],
                       c.geom.plo[2], 1. / c.geom.dx[0], 1. / c.geom.dx[1], 1. / c.geom.dx[2], c.ncell[0],
                       c.ncell[1], c.ncell[2], dpos, drd, did, dout);
  };
  switch (c.nf) {
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 6: run(std::integral_constant<int, 6>{}); break;
    default: set_last_error("unsupported fuel count"); st = PELEMP_E_BAD_CONFIG;
  }
  int outside = 0;
  if (!st) {
    PM_HIP(hipGetLastError());
    PM_HIP(hipMemcpy(&outside, dout, 4, hipMemcpyDeviceToHost));
    if (outside) {
      set_last_error("Bad particle");
      st = PELEMP_E_BAD_INJECTION;
    } else {
      st = append_staged(c, n, dpos, drd, reinterpret_cast<const int64_t*>(did), next_id + n);
    }
  }
  (void)hipFree(dpos);
  (void)hipFree(drd);
  (void)hipFree(did);
  (void)hipFree(dY);
  (void)hipFree(dout);
  return st;
}

int pelemp_spray_write_injection_log(const char* path, int32_t njets, const char* const* names,
                                     const pelemp_spray_jet* jets) {
  if (!path || njets < 0 || (njets > 0 && (!names || !jets))) return PELEMP_E_ARG;
  std::ofstream file(path, std::ios::out | std::ios::trunc);
  if (!file.good()) {
    set_last_error(std::string("cannot open ") + path);
    return PELEMP_E_ARG;
  }
  file << njets << "\n";
  file.precision(15);
  for (int32_t i = 0; i < njets; ++i) {
    const pelemp_spray_jet& j = jets[i];
    file << names[i] << " " << j.num_ppp << " " << j.sum_inj_mass << " " << j.sum_inj_time << " " << j.min_parcel
         << " " << j.total_inj_mass << " " << j.total_inj_time << "\n";
  }
  file.close();
  if (!file.good()) {
    set_last_error("Problem writing injection file");
    return PELEMP_E_ARG;
  }
  return PELEMP_OK;
}

int pelemp_spray_read_injection_log(const char* path, int32_t njets, const char* const* names,
                                    pelemp_spray_jet* jets) {
  if (!path || njets < 0 || (njets > 0 && (!names || !jets))) return PELEMP_E_ARG;
  if (njets == 0) return PELEMP_OK;
  // PostInitRestart's name checks (SprayIO.cpp:105-118): names must be set;
  // repeated names next to each other are an error (std::unique)
  for (int32_t i = 0; i < njets; ++i) {
    if (!names[i] || names[i][0] == '\0') {
      set_last_error("Jet not provided name");
      return PELEMP_E_ARG;
    }
    if (i > 0 && std::strcmp(names[i], names[i - 1]) == 0) {
      set_last_error("Duplicate jet names detected");
      return PELEMP_E_ARG;
    }
  }
  std::ifstream in(path);
  if (!in.good()) return PELEMP_OK;  // no file: jets keep their state
  int in_numjets = 0;
  in >> in_numjets;
  for (int i = 0; i < in_numjets; ++i) {
    std::string name;
    double v[6];
    in >> name >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5];
    if (in.fail()) {
      set_last_error("malformed injection file");
      return PELEMP_E_ARG;
    }
    for (int32_t mj = 0; mj < njets; ++mj) {
      if (name == names[mj]) {
        jets[mj].num_ppp = v[0];
        jets[mj].sum_inj_mass = v[1];
        jets[mj].sum_inj_time = v[2];
        jets[mj].min_parcel = v[3];
        jets[mj].total_inj_mass = v[4];
        jets[mj].total_inj_time = v[5];
      }
    }
  }
  return PELEMP_OK;
}

// all particles in the caller's units (AoS meaning)
static int fetch_all(pelemp_spray_ctx* h, std::vector<double>& pos, std::vector<double>& rd,
                     std::vector<int64_t>& ids) {
  int64_t n = 0;
  int st = pelemp_spray_num_particles(h, &n);
  if (st) return st;
  const int nsr = 10 + h->c.nf;
  pos.resize(3 * std::max<int64_t>(n, 1));
  rd.resize(nsr * std::max<int64_t>(n, 1));
  ids.resize(std::max<int64_t>(n, 1));
  st = pelemp_spray_get_particles(h, &n, pos.data(), rd.data(), ids.data());
  pos.resize(3 * n);
  rd.resize(nsr * n);
  ids.resize(n);
  return st;
}

int pelemp_spray_write_ascii(pelemp_spray_ctx* h, const char* path) {
  if (!h || !path) return PELEMP_E_ARG;
  std::vector<double> pos, rd;
  std::vector<int64_t> ids;
  int st = fetch_all(h, pos, rd, ids);
  if (st) return st;
  const int nsr = 10 + h->c.nf;
  std::ofstream file(path, std::ios::out | std::ios::trunc);
  if (!file.good()) {
    set_last_error(std::string("cannot open ") + path);
    return PELEMP_E_ARG;
  }
  file << ids.size() << '\n';
  file.precision(15);
  for (size_t i = 0; i < ids.size(); ++i) {
    for (int d = 0; d < 3; ++d) file << pos[3 * i + d] << ' ';
    file << io::amrex_id(ids[i]) << ' ' << io::amrex_cpu(ids[i]) << ' ';
    for (int q = 0; q < nsr; ++q) file << rd[nsr * i + q] << ' ';
    file << '\n';
  }
  file.close();
  if (!file.good()) {
    set_last_error("Problem writing particle file");
    return PELEMP_E_ARG;
  }
  return PELEMP_OK;
}

int pelemp_spray_checkpoint(pelemp_spray_ctx* h, const char* dir, int32_t nfuel_names,
                            const char* const* fuel_names) {
  if (!h || !dir) return PELEMP_E_ARG;
  Ctx& c = h->c;
  std::vector<double> pos, rd;
  std::vector<int64_t> ids;
  int st = fetch_all(h, pos, rd, ids);
  if (st) return st;
  int64_t next_id = 0, next_slot = 0;
  st = read_counters(c, &next_id, &next_slot);
  if (st) return st;
  io::Box dom;
  for (int d = 0; d < 3; ++d) {
    dom.lo[d] = c.geom.domain_lo[d];
    dom.hi[d] = c.geom.domain_hi[d];
  }
  std::string err;
  st = io::write_checkpoint(dir, "particles", static_cast<int64_t>(ids.size()), pos.data(), rd.data(), ids.data(),
                            10 + c.nf, spray_comp_names(c, nfuel_names, fuel_names), next_id, {dom}, c.geom.plo,
                            c.geom.dx, err);
  if (st) set_last_error(err);
  return st;
}

int pelemp_spray_restart(pelemp_spray_ctx* h, const char* dir) {
  if (!h || !dir) return PELEMP_E_ARG;
  Ctx& c = h->c;
  const std::string pdir = std::string(dir) + "/particles";
  io::HeaderInfo hi;
  std::string err;
  int st = io::read_header(pdir, hi, err);
  if (st) {
    set_last_error(err);
    return st;
  }
  if (hi.nreal != 10 + c.nf || hi.nint != 0) {
    set_last_error("checkpoint has " + std::to_string(hi.nreal) + " real and " + std::to_string(hi.nint) +
                   " int components; the context's droplets have " + std::to_string(10 + c.nf) + " and 0");
    return PELEMP_E_BAD_CONFIG;
  }
  std::vector<double> pos, rd;
  std::vector<int64_t> ids;
  st = io::read_level(pdir, hi, 0, pos, rd, ids, err);
  if (st) {
    set_last_error(err);
    return st;
  }
  const int64_t n = static_cast<int64_t>(ids.size());
  st = pelemp_spray_set_particles(h, n, pos.data(), rd.data(), ids.data());
  if (st) return st;
  int64_t cur_id = 0, next_slot = 0;
  st = read_counters(c, &cur_id, &next_slot);
  if (st) return st;
  return write_counters(c, std::max(cur_id, hi.next_id), next_slot);
}

int pelemp_particles_write_checkpoint(const char* dir, const char* name, int64_t n, const double* pos,
                                      const double* rdata, const int64_t* ids, int32_t nreal,
                                      const char* const* real_names, int64_t next_id, int32_t nboxes,
                                      const int32_t* box_lo, const int32_t* box_hi, const double plo[3],
                                      const double dx[3]) {
  if (!dir || !name || nreal < 0 || (nreal > 0 && !real_names) || nboxes < 1 || !box_lo || !box_hi ||
      (nboxes > 1 && (!plo || !dx)))
    return PELEMP_E_ARG;
  std::vector<std::string> names;
  for (int q = 0; q < nreal; ++q) names.push_back(real_names[q] ? real_names[q] : "");
  std::vector<io::Box> boxes(nboxes);
  for (int b = 0; b < nboxes; ++b)
    for (int d = 0; d < 3; ++d) {
      boxes[b].lo[d] = box_lo[3 * b + d];
      boxes[b].hi[d] = box_hi[3 * b + d];
    }
  const double zero[3] = {0.0, 0.0, 0.0}, one[3] = {1.0, 1.0, 1.0};
  std::string err;
  const int st = io::write_checkpoint(dir, name, n, pos, rdata, ids, nreal, names, next_id, boxes, plo ? plo : zero,
                                      dx ? dx : one, err);
  if (st) set_last_error(err);
  return st;
}

int pelemp_particles_checkpoint_info(const char* dir, const char* name, int32_t level, int64_t* n,
                                     int32_t* nreal, int64_t* next_id) {
  if (!dir || !name || !n || !nreal || !next_id) return PELEMP_E_ARG;
  io::HeaderInfo hi;
  std::string err;
  int st = io::read_header(std::string(dir) + "/" + name, hi, err);
  if (!st && (level < 0 || level > hi.finest)) {
    err = "particle checkpoint: level " + std::to_string(level) + " is not in the checkpoint";
    st = PELEMP_E_ARG;
  }
  if (st) {
    set_last_error(err);
    return st;
  }
  *n = io::level_count(hi, level);
  *nreal = hi.nreal;
  *next_id = hi.next_id;
  return PELEMP_OK;
}

int pelemp_particles_read_checkpoint(const char* dir, const char* name, int32_t level, int64_t capacity,
                                     int64_t* n, double* pos, double* rdata, int64_t* ids) {
  if (!dir || !name || !n || capacity < 0) return PELEMP_E_ARG;
  const std::string pdir = std::string(dir) + "/" + name;
  io::HeaderInfo hi;
  std::string err;
  std::vector<double> p, r;
  std::vector<int64_t> id;
  int st = io::read_header(pdir, hi, err);
  if (!st) st = io::read_level(pdir, hi, level, p, r, id, err);
  if (st) {
    set_last_error(err);
    return st;
  }
  const int64_t m = static_cast<int64_t>(id.size());
  if (m > capacity) {
    set_last_error("particle checkpoint: " + std::to_string(m) + " particles exceed the capacity " +
                   std::to_string(capacity));
    return PELEMP_E_ARG;
  }
  if (m > 0 && (!pos || !ids || (hi.nreal > 0 && !rdata))) return PELEMP_E_ARG;
  if (m > 0) {
    std::memcpy(pos, p.data(), sizeof(double) * p.size());
    if (hi.nreal > 0) std::memcpy(rdata, r.data(), sizeof(double) * r.size());
    std::memcpy(ids, id.data(), sizeof(int64_t) * id.size());
  }
  *n = m;
  return PELEMP_OK;
}

int pelemp_spray_set_eb(pelemp_spray_ctx* h, const pelemp_eb_view* eb) {
  if (!h) return PELEMP_E_ARG;
  Ctx& c = h->c;
  if (!eb) {
    c.eb_set = false;
    c.eb_on = false;
    c.n_eb_cells = 0;
    return PELEMP_OK;
  }
  if (c.form != PELEMP_FORM_FP64) {
    set_last_error("embedded boundaries need the FP64 form");
    return PELEMP_E_BAD_CONFIG;
  }
  if (!eb->flags || !eb->vfrac || !eb->ccent || !eb->bcent || !eb->bnorm) return PELEMP_E_ARG;
  for (int d = 0; d < 3; ++d)
    if (eb->lo[d] > c.geom.domain_lo[d] - 1 || eb->hi[d] < c.geom.domain_hi[d] + 1) {
      set_last_error("the EB box must cover the domain grown by one cell");
      return PELEMP_E_ARG;
    }
  if (!c.branches) {
    // the EB runs in the branch kernel: switch an empty context to its
    // layout (breakup state in the hot fields, wider slot arrays)
    int64_t n = 0;
    int st = pelemp_spray_num_particles(h, &n);
    if (st) return st;
    if (n > 0 || c.nslots > 0) {
      set_last_error("attach the EB before handing over droplets");
      return PELEMP_E_BAD_CONFIG;
    }
    (void)hipFree(c.fbuf[0]);
    (void)hipFree(c.fbuf[1]);
    (void)hipFree(c.cold_id);
    c.fbuf[0] = c.fbuf[1] = nullptr;
    c.cold_id = nullptr;
    c.branches = true;
    c.hot_brk = true;
    c.nhot = 13 + c.nf;
    c.slot_cap = 4 * c.cap;
    const size_t fb = c.wsize * c.nhot * static_cast<size_t>(c.cap);
    st = alloc(c, &c.fbuf[0], fb);
    st |= alloc(c, &c.fbuf[1], fb);
    st |= alloc(c, reinterpret_cast<void**>(&c.cold_id), 8 * c.slot_cap);
    st |= alloc_branch_buffers(c);
    if (st) return PELEMP_E_NOMEM;
    const int64_t ids[4] = {c.next_id, 0, 0, 0};
    PM_HIP(hipMemcpy(c.sb_ids, ids, sizeof(ids), hipMemcpyHostToDevice));
  }
  // cut or covered cells, and the domain cells next to them
  const int nb[3] = {eb->hi[0] - eb->lo[0] + 1, eb->hi[1] - eb->lo[1] + 1, eb->hi[2] - eb->lo[2] + 1};
  const size_t ncb = static_cast<size_t>(nb[0]) * nb[1] * nb[2];
  std::vector<uint32_t> hf(ncb);
  PM_HIP(hipMemcpy(hf.data(), eb->flags, 4 * ncb, hipMemcpyDeviceToHost));
  auto irregular = [&](int i, int j, int k) {  // absolute indices
    if (i < eb->lo[0] || i > eb->hi[0] || j < eb->lo[1] || j > eb->hi[1] || k < eb->lo[2] || k > eb->hi[2])
      return false;
    return (hf[(static_cast<size_t>(k - eb->lo[2]) * nb[1] + (j - eb->lo[1])) * nb[0] + (i - eb->lo[0])] & 3u) != 0u;
  };
  bool any = false;
  for (size_t i = 0; i < ncb && !any; ++i) any = (hf[i] & 3u) != 0u;
  std::vector<int32_t> cells;
  if (any) {
    // deposit targets: cells of the domain grown by one next to the EB
    const int* dl = c.geom.domain_lo;
    for (int k = -1; k <= c.ncell[2]; ++k)
      for (int j = -1; j <= c.ncell[1]; ++j)
        for (int i = -1; i <= c.ncell[0]; ++i) {
          bool near = false;
          for (int dz = -1; dz <= 1 && !near; ++dz)
            for (int dy = -1; dy <= 1 && !near; ++dy)
              for (int dx = -1; dx <= 1 && !near; ++dx)
                near = irregular(dl[0] + i + dx, dl[1] + j + dy, dl[2] + k + dz);
          if (near) cells.push_back(static_cast<int32_t>(eb_cell_code(i, j, k, c.ncell)));
        }
  }
  if (c.eb_cells) {
    (void)hipFree(c.eb_cells);
    c.eb_cells = nullptr;
  }
  if (!cells.empty()) {
    if (alloc(c, reinterpret_cast<void**>(&c.eb_cells), 4 * cells.size())) return PELEMP_E_NOMEM;
    PM_HIP(hipMemcpy(c.eb_cells, cells.data(), 4 * cells.size(), hipMemcpyHostToDevice));
  }
  if (any && !c.eb_rkey) {
    int st = alloc(c, reinterpret_cast<void**>(&c.eb_rkey), 4 * static_cast<size_t>(c.cap));
    st |= alloc(c, reinterpret_cast<void**>(&c.eb_rval), 8 * static_cast<size_t>(5 + c.nf) * c.cap);
    if (st) return PELEMP_E_NOMEM;
  }
  c.eb = *eb;
  c.eb_set = true;
  c.eb_on = any;
  c.n_eb_cells = static_cast<int32_t>(cells.size());
  return PELEMP_OK;
}

}  // extern "C"

namespace pelemp {
// updateParticles for a state in FP64 (state->data) or FP32 (state32, with
// the box and components in `state`).
static int update_impl(pelemp_spray_ctx* h, const pelemp_fab_view* state, const float* state32,
                       pelemp_fab_view* source, double flow_dt, int32_t state_ghosts, int32_t source_ghosts,
                       int32_t isVirt, int32_t isGhost, int32_t do_move, double spray_cfl_lev, void* stream) {
  if (!h || !state || !source || !(state->data || state32) || !source->data) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  Ctx& c = h->c;
  if (state32 && c.pelec) {
    set_last_error("an FP32 gas state is supported in the PeleLMeX flavour only");
    return PELEMP_E_BAD_CONFIG;
  }
  const pelemp_spray_comps& S = c.comps;
  const int st_need = std::max({S.rhoIndx, S.momIndx + 2, S.specIndx + c.ns - 1, S.utempIndx});
  const int sr_need = std::max({S.rhoSrcIndx, S.momSrcIndx + 2, S.specSrcIndx + c.nf - 1, S.engSrcIndx});
  if (state->ncomp <= st_need || source->ncomp <= sr_need) {
    set_last_error("state or source FAB has too few components for SprayComps");
    return PELEMP_E_ARG;
  }
  if (!(flow_dt > 0.)) {
    set_last_error("flow_dt must be positive");
    return PELEMP_E_ARG;
  }
  if (c.form == PELEMP_FORM_SURROGATE && !c.has_surr) {
    set_last_error("surrogate form: no network installed (pelemp_spray_set_surrogate)");
    return PELEMP_E_BAD_CONFIG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // particle sub-cycles (SprayParticles.cpp:221-229)
  const double sub_cfl = 0.5;
  double sub_dt = flow_dt;
  int num_iter = 1;
  if (do_move && spray_cfl_lev > sub_cfl) {
    num_iter = static_cast<int>(std::ceil(spray_cfl_lev / sub_cfl));
    sub_dt = flow_dt / static_cast<double>(num_iter);
  }
  const double vol = c.geom.dx[0] * c.geom.dx[1] * c.geom.dx[2];
  const double inv_vol = 1. / vol;
  PassParams p;
  p.flow_dt = flow_dt;
  p.sub_dt = sub_dt;
  p.cur_coef = -inv_vol * sub_dt / flow_dt;
  p.do_move = do_move ? 1 : 0;
  p.is_ghost = isGhost ? 1 : 0;
  // the state box and the source box: the level's box grown by the ghost
  // widths (SprayParticles.cpp:250-251), within the FABs handed in
  p.state_ghosts = state_ghosts;
  p.source_ghosts = source_ghosts;
  p.at_bounds = 0;
  for (int d = 0; d < 3; ++d)
    if (c.geom.bndry_lo[d] != 0 || c.geom.bndry_hi[d] != 0) p.at_bounds = 1;
  p.state = state;
  p.state32 = state32;
  p.source = source;
  // branch flags (SprayParticles.cpp:162-164,289-299); the splash model
  // runs where walls are (tile_at_bndry, WallFunctions.H:7-35)
  const bool isActive = !(isVirt || isGhost);
  const bool do_splash = c.sd.do_splash != 0 && isActive && do_move;
  p.branches = c.branches ? 1 : 0;
  p.do_splash_box = (do_splash && (p.at_bounds || c.eb_on)) ? 1 : 0;
  p.make_new_drops = ((c.sd.do_breakup > 0 || p.do_splash_box) && isActive && do_move) ? 1 : 0;
  if (c.branches && num_iter > 1) {
    set_last_error("breakup / splash / films need a single particle sub-cycle (spray_cfl_lev <= 0.5)");
    return PELEMP_E_BAD_CONFIG;
  }
  if (c.pelec) {
    const int stc = launch_pelec_convert(c, *state, s);
    if (stc != PELEMP_OK) return stc;
  }
  for (int it = 0; it < num_iter; ++it) {
    // overwrite mode: the first sub-cycle writes the source, later ones add
    p.src_overwrite = (c.src_mode == PELEMP_SOURCE_OVERWRITE && it == 0) ? 1 : 0;
    int32_t* log_slot = (c.npass < Ctx::kCountLog) ? c.count_log + c.npass : nullptr;
    ++c.npass;
    // the last sort cleared the pass counters, unless a pass stopped early
    if (c.counters_dirty) PM_HIP(hipMemsetAsync(c.mov_lo, 0, 3 * 4 * static_cast<size_t>(c.ncells), s));
    c.counters_dirty = true;
    if (p.do_splash_box) {
      hipEvent_t evf;
      ctx_time_begin(c, 2, s, &evf);
      const int st = run_film_fab(c, s);
      if (st != PELEMP_OK) return st;
      ctx_time_end(c, 2, s, evf);
    }
    hipEvent_t ev0;
    ctx_time_begin(c, 0, s, &ev0);
    const int st = (c.form == PELEMP_FORM_FP64)   ? launch_update_fp64(c, p, s)
                   : (c.form == PELEMP_FORM_FP32) ? launch_update_fp32(c, p, s)
                                                  : launch_update_surr(c, p, s);
    if (st != PELEMP_OK) return st;
    ctx_time_end(c, 0, s, ev0);
    if (c.eb_on && c.n_eb_cells > 0) {
      const int ste = run_eb_redirect(c, p, s);
      if (ste != PELEMP_OK) return ste;
    }
    if (p.make_new_drops) {
      hipEvent_t evc;
      ctx_time_begin(c, 3, s, &evc);
      const int stc = run_children(c, p, s);
      if (stc != PELEMP_OK) return stc;
      ctx_time_end(c, 3, s, evc);
    }
    c.cur = 1 - c.cur;
    const int st2 = run_sort_incremental(c, s, p.make_new_drops ? c.n_new : nullptr, log_slot);
    if (st2 != PELEMP_OK) return st2;
  }
  return PELEMP_OK;
}

}  // namespace pelemp

extern "C" {

int pelemp_spray_update(pelemp_spray_ctx* h, const pelemp_fab_view* state, pelemp_fab_view* source,
                        double flow_dt, double time, int32_t state_ghosts, int32_t source_ghosts, int32_t isVirt,
                        int32_t isGhost, int32_t do_move, double spray_cfl_lev, void* stream) {
  (void)time;
  if (state && !state->data) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  return update_impl(h, state, nullptr, source, flow_dt, state_ghosts, source_ghosts, isVirt, isGhost, do_move,
                     spray_cfl_lev, stream);
}

int pelemp_spray_update_f32gas(pelemp_spray_ctx* h, const pelemp_fab_view_f32* state, pelemp_fab_view* source,
                               double flow_dt, double time, int32_t state_ghosts, int32_t source_ghosts,
                               int32_t isVirt, int32_t isGhost, int32_t do_move, double spray_cfl_lev,
                               void* stream) {
  (void)time;
  if (!state || !state->data) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  pelemp_fab_view box{};
  box.data = nullptr;
  for (int d = 0; d < 3; ++d) {
    box.lo[d] = state->lo[d];
    box.hi[d] = state->hi[d];
  }
  box.ncomp = state->ncomp;
  return update_impl(h, &box, state->data, source, flow_dt, state_ghosts, source_ghosts, isVirt, isGhost, do_move,
                     spray_cfl_lev, stream);
}

int pelemp_spray_move_kick_drift(pelemp_spray_ctx* h, const pelemp_fab_view* state, pelemp_fab_view* source,
                                 double dt, double time, int32_t isVirt, int32_t isGhost, int32_t state_ghosts,
                                 int32_t source_ghosts, int32_t do_move, double spray_cfl_lev, void* stream) {
  return pelemp_spray_update(h, state, source, dt, time, state_ghosts, source_ghosts, isVirt, isGhost, do_move,
                             spray_cfl_lev, stream);
}

int pelemp_spray_set_source_mode(pelemp_spray_ctx* h, int32_t mode) {
  if (!h || (mode != PELEMP_SOURCE_ACCUMULATE && mode != PELEMP_SOURCE_OVERWRITE)) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  h->c.src_mode = mode;
  return PELEMP_OK;
}

int pelemp_spray_set_params(pelemp_spray_ctx* h, const pelemp_spray_params* p) {
  if (!h || !p) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  // SpraySetup.cpp:144-160
  if (!(p->max_parcel_size > 0.) || p->breakup_parcel_factor < 0. || p->breakup_parcel_factor > 1.) {
    set_last_error("max_parcel_size must be positive and breakup_parcel_factor in [0, 1]");
    return PELEMP_E_BAD_CONFIG;
  }
  h->c.params = *p;
  if (h->c.pelec) h->c.params.avg_inject_mass *= cgs::kVol;
  return PELEMP_OK;
}

int pelemp_spray_get_params(const pelemp_spray_ctx* h, pelemp_spray_params* out) {
  if (!h || !out) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  *out = h->c.params;
  if (h->c.pelec) out->avg_inject_mass /= cgs::kVol;
  return PELEMP_OK;
}

int pelemp_spray_set_surrogate(pelemp_spray_ctx* h, const float* params_host, const pelemp_surrogate_norm* norm) {
  if (!h || !params_host || !norm) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  Ctx& c = h->c;
  if (c.form != PELEMP_FORM_SURROGATE) {
    set_last_error("the context was not created with PELEMP_FORM_SURROGATE");
    return PELEMP_E_BAD_CONFIG;
  }
  for (int q = 0; q < kSurrOut; ++q)
    if (!(norm->out_std[q] > 0.f)) {
      set_last_error("surrogate norm: out_std must be positive");
      return PELEMP_E_ARG;
    }
  std::vector<SurrWeights> w(1);
  surr_pack(params_host, w[0]);
  PM_HIP(hipDeviceSynchronize());  // no update may be reading the old network
  PM_HIP(hipMemcpy(c.surr_w, w.data(), sizeof(SurrWeights), hipMemcpyHostToDevice));
  c.surr_norm = *norm;
  c.has_surr = true;
  return PELEMP_OK;
}

int pelemp_spray_surrogate_counts(pelemp_spray_ctx* h, int64_t* n_surrogate, int64_t* n_fallback) {
  if (!h || !n_surrogate || !n_fallback) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  Ctx& c = h->c;
  if (c.form != PELEMP_FORM_SURROGATE) {
    *n_surrogate = 0;
    *n_fallback = 0;
    return PELEMP_OK;
  }
  PM_HIP(hipDeviceSynchronize());
  unsigned long long v[2] = {0, 0};
  PM_HIP(hipMemcpy(v, c.surr_cnt, sizeof(v), hipMemcpyDeviceToHost));
  PM_HIP(hipMemset(c.surr_cnt, 0, sizeof(v)));
  *n_surrogate = static_cast<int64_t>(v[0]);
  *n_fallback = static_cast<int64_t>(v[1]);
  return PELEMP_OK;
}

// Waits for the stream and returns the sticky device error (0 if none).
int pelemp_spray_sync(pelemp_spray_ctx* h, void* stream) {
  if (!h) return PELEMP_E_ARG;
  PM_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  int32_t e = 0;
  PM_HIP(hipMemcpy(&e, h->c.err, 4, hipMemcpyDeviceToHost));
  if (e != 0) {
    set_last_error("device-side check failed (code " + std::to_string(e) + ")");
    PM_HIP(hipMemset(h->c.err, 0, 4));
  }
  return e;
}

int pelemp_spray_est_timestep(pelemp_spray_ctx* h, double cfl, double max_jet_vel, double* dt, void* stream) {
  if (!h || !dt) return PELEMP_E_ARG;
  Ctx& c = h->c;
  if (c.sd.fixed_parts) {
    *dt = -1.;
    return PELEMP_OK;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned long long init = 0x7fefffffffffffffULL;  // DBL_MAX bits
  PM_HIP(hipMemcpyAsync(c.dscratch, &init, 8, hipMemcpyHostToDevice, s));
  const double dxi[3] = {1. / c.geom.dx[0], 1. / c.geom.dx[1], 1. / c.geom.dx[2]};
  auto* out = reinterpret_cast<unsigned long long*>(c.dscratch);
  if (c.form == PELEMP_FORM_FP64) {
    hipLaunchKernelGGL((est_dt_kernel<double>), dim3(1024), dim3(256), 0, s, c.cstart + c.ncells, c.perm,
                       static_cast<const double*>(c.fbuf[c.cur]), c.cap, dxi[0], dxi[1], dxi[2], cfl, out);
  } else {
    hipLaunchKernelGGL((est_dt_kernel<float>), dim3(1024), dim3(256), 0, s, c.cstart + c.ncells, c.perm,
                       static_cast<const float*>(c.fbuf[c.cur]), c.cap, dxi[0], dxi[1], dxi[2], cfl, out);
  }
  PM_HIP(hipGetLastError());
  unsigned long long bits = 0;
  PM_HIP(hipMemcpyAsync(&bits, c.dscratch, 8, hipMemcpyDeviceToHost, s));
  PM_HIP(hipStreamSynchronize(s));
  double v;
  memcpy(&v, &bits, 8);
  // no droplets: the reduction's identity, the largest double (the
  // reference starts from std::numeric_limits<Real>::max())
  int32_t t = 0;
  PM_HIP(hipMemcpy(&t, c.cstart + c.ncells, 4, hipMemcpyDeviceToHost));
  double d = (t == 0) ? std::numeric_limits<double>::max() : v;
  // injection bound of the jets (SprayParticles.cpp:125-131)
  if (c.pelec) max_jet_vel *= cgs::kLen;
  if (max_jet_vel > 0.) d = std::min(d, cfl * c.geom.dx[0] / max_jet_vel);
  *dt = d;
  return PELEMP_OK;
}

int pelemp_spray_derive_count(const pelemp_spray_ctx* h, int32_t derive_plot_species, int32_t* n) {
  if (!h || !n) return PELEMP_E_ARG;
  *n = kDeriveBase + ((derive_plot_species == 1 && h->c.nf > 1) ? h->c.nf : 0);
  return PELEMP_OK;
}

int pelemp_spray_compute_derived(pelemp_spray_ctx* h, pelemp_fab_view* mf_var, int32_t start_indx,
                                 int32_t derive_plot_species, void* stream) {
  if (!h || !mf_var || !mf_var->data) {
    set_last_error("compute_derived: invalid argument");
    return PELEMP_E_ARG;
  }
  Ctx& c = h->c;
  const int nvar = kDeriveBase + ((derive_plot_species == 1 && c.nf > 1) ? c.nf : 0);
  if (start_indx < 0 || start_indx + nvar > mf_var->ncomp) {
    set_last_error("compute_derived: the FAB has fewer components than start_indx + the spray variables");
    return PELEMP_E_ARG;
  }
  for (int d = 0; d < 3; ++d) {
    if (mf_var->lo[d] > c.geom.domain_lo[d] || mf_var->hi[d] < c.geom.domain_hi[d]) {
      set_last_error("compute_derived: the FAB must cover the domain");
      return PELEMP_E_ARG;
    }
  }
  XferFab mf;
  mf.data = mf_var->data;
  for (int d = 0; d < 3; ++d) {
    mf.lo[d] = mf_var->lo[d];
    mf.n[d] = mf_var->hi[d] - mf_var->lo[d] + 1;
  }
  mf.cstride = static_cast<long long>(mf.n[0]) * mf.n[1] * mf.n[2];
  DeriveGeom g;
  for (int d = 0; d < 3; ++d) {
    g.dlo[d] = c.geom.domain_lo[d];
    g.ncell[d] = c.ncell[d];
    g.ntile[d] = c.ntile[d];
  }
  g.cell_vol = c.geom.dx[0] * c.geom.dx[1] * c.geom.dx[2];
  g.face_area = 1. * c.geom.dx[0] * c.geom.dx[0];  // SprayDerive.cpp:85
  g.pelec = c.pelec ? 1 : 0;
  for (int q = 0; q < kDeriveBase + PELEMP_MAX_FUEL; ++q) g.scale[q] = 1.;
  if (c.pelec) cgs::derived_scale(c.nf, g.scale);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned grid = static_cast<unsigned>(std::min<long long>((mf.cstride + 255) / 256, 16384));
  auto run = [&](auto nft) {
    constexpr int NF = decltype(nft)::value;
    const FuelConst<double, NF> fc = make_fuel_const<double, NF>(c.sd, c.mw.data());
    if (c.form == PELEMP_FORM_FP64) {
      hipLaunchKernelGGL((derive_kernel<NF, double>), dim3(grid), dim3(256), 0, s, c.cstart, c.perm,
                         static_cast<const double*>(c.fbuf[c.cur]), nullptr, c.cap, fc, g, mf, start_indx, nvar);
    } else {
      hipLaunchKernelGGL((derive_kernel<NF, float>), dim3(grid), dim3(256), 0, s, c.cstart, c.perm,
                         static_cast<const float*>(c.fbuf[c.cur]), c.mbuf[c.cur], c.cap, fc, g, mf, start_indx,
                         nvar);
    }
  };
  switch (c.nf) {
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 6: run(std::integral_constant<int, 6>{}); break;
    default: set_last_error("unsupported fuel count"); return PELEMP_E_BAD_CONFIG;
  }
  PM_HIP(hipGetLastError());
  return PELEMP_OK;
}

int pelemp_spray_transfer_source(pelemp_spray_ctx* h, int32_t source_ghosts, pelemp_fab_view* tmp,
                                 pelemp_fab_view* act, void* stream) {
  if (!h || !tmp || !act || !tmp->data || !act->data || tmp->data == act->data || source_ghosts < 0) {
    set_last_error("transfer_source: invalid argument");
    return PELEMP_E_ARG;
  }
  Ctx& c = h->c;
  const pelemp_geom& g = c.geom;
  int ng_tmp = 1 << 30, ng_act = 1 << 30;
  for (int d = 0; d < 3; ++d) {
    ng_tmp = std::min(ng_tmp, std::min(g.domain_lo[d] - tmp->lo[d], tmp->hi[d] - g.domain_hi[d]));
    ng_act = std::min(ng_act, std::min(g.domain_lo[d] - act->lo[d], act->hi[d] - g.domain_hi[d]));
  }
  if (ng_tmp < 0 || ng_act < 0) {
    set_last_error("transfer_source: the FABs must cover the domain");
    return PELEMP_E_ARG;
  }
  XferMap map{};
  if (tmp->ncomp == act->ncomp) {  // MultiFab::Add of every component
    if (tmp->ncomp > kXferMaxPairs) {
      set_last_error("transfer_source: too many components");
      return PELEMP_E_ARG;
    }
    for (int q = 0; q < tmp->ncomp; ++q) {
      map.from[q] = q;
      map.to[q] = q;
    }
    map.npair = tmp->ncomp;
  } else {  // addSpraySrc (SprayParticles.H:299-319)
    const pelemp_spray_comps& S = c.comps;
    int n = 0;
    for (int dir = 0; dir < 3; ++dir) {
      map.from[n] = S.momSrcIndx + dir;
      map.to[n++] = S.momIndx + dir;
    }
    map.from[n] = S.rhoSrcIndx;
    map.to[n++] = S.rhoIndx;
    map.from[n] = S.engSrcIndx;
    map.to[n++] = S.engIndx;
    for (int sp = 0; sp < c.nf; ++sp) {
      map.from[n] = S.specSrcIndx + sp;
      map.to[n++] = S.specIndx + c.sd.dep_indx[sp];
    }
    map.npair = n;
    for (int q = 0; q < n; ++q) {
      if (map.from[q] < 0 || map.from[q] >= tmp->ncomp || map.to[q] < 0 || map.to[q] >= act->ncomp) {
        set_last_error("transfer_source: spray component outside the FABs");
        return PELEMP_E_ARG;
      }
    }
  }
  auto fab = [](const pelemp_fab_view& v) {
    XferFab f;
    f.data = v.data;
    for (int d = 0; d < 3; ++d) {
      f.lo[d] = v.lo[d];
      f.n[d] = v.hi[d] - v.lo[d] + 1;
    }
    f.cstride = static_cast<long long>(f.n[0]) * f.n[1] * f.n[2];
    return f;
  };
  XferGeom xg;
  xg.nghost = std::min(std::min(ng_act, static_cast<int>(source_ghosts)), ng_tmp);
  for (int d = 0; d < 3; ++d) {
    xg.dlo[d] = g.domain_lo[d];
    xg.N[d] = g.domain_hi[d] - g.domain_lo[d] + 1;
    xg.periodic[d] = (g.bndry_lo[d] == 0) ? 1 : 0;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const XferFab ft = fab(*tmp), fa = fab(*act);
  long long ncell = 1;
  for (int d = 0; d < 3; ++d) ncell *= xg.N[d] + 2 * xg.nghost;
  const unsigned grid = static_cast<unsigned>(std::min<long long>((ncell + 255) / 256, 16384));
  hipLaunchKernelGGL(transfer_add_kernel, dim3(grid), dim3(256), 0, s, ft, fa, map, xg);
  const long long ntmp = ft.cstride * tmp->ncomp;
  const unsigned gz = static_cast<unsigned>(std::min<long long>((ntmp + 255) / 256, 16384));
  hipLaunchKernelGGL(zero_fill_kernel, dim3(gz), dim3(256), 0, s, tmp->data, ntmp);
  PM_HIP(hipGetLastError());
  return PELEMP_OK;
}

int pelemp_spray_pass_counts(pelemp_spray_ctx* h, int64_t* out, int32_t max, int32_t* n_out) {
  if (!h || !n_out || (max > 0 && !out)) return PELEMP_E_ARG;
  Ctx& c = h->c;
  PM_HIP(hipDeviceSynchronize());
  const int64_t m = std::min<int64_t>(std::min<int64_t>(c.npass, Ctx::kCountLog), max);
  std::vector<int32_t> tmp(static_cast<size_t>(std::max<int64_t>(m, 1)));
  if (m > 0) PM_HIP(hipMemcpy(tmp.data(), c.count_log, 4 * m, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < m; ++i) out[i] = tmp[i];
  *n_out = static_cast<int32_t>(m);
  c.npass = 0;
  return PELEMP_OK;
}

int pelemp_spray_set_timing(pelemp_spray_ctx* h, int32_t enable) {
  if (!h) return PELEMP_E_ARG;
  h->c.timing = enable != 0;
  harvest_timing(h->c);
  for (int i = 0; i < 4; ++i) {
    h->c.t_ms[i] = 0;
    h->c.t_n[i] = 0;
  }
  return PELEMP_OK;
}

int pelemp_spray_get_timing(pelemp_spray_ctx* h, double* ms, int64_t* launches, int32_t nslots) {
  if (!h) return PELEMP_E_ARG;
  harvest_timing(h->c);
  for (int i = 0; i < nslots && i < 4; ++i) {
    if (ms) ms[i] = h->c.t_ms[i];
    if (launches) launches[i] = h->c.t_n[i];
  }
  return PELEMP_OK;
}

int pelemp_spray_device_bytes(const pelemp_spray_ctx* h, int64_t* bytes) {
  if (!h || !bytes) return PELEMP_E_ARG;
  *bytes = static_cast<int64_t>(h->c.bytes);
  return PELEMP_OK;
}

int pelemp_math_eval(int32_t fn, int64_t n, const double* x, const double* y, double* out) {
  if (n <= 0) return PELEMP_OK;
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  PM_HIP(hipMalloc(&dx, 8 * n));
  PM_HIP(hipMalloc(&dout, 8 * n));
  PM_HIP(hipMemcpy(dx, x, 8 * n, hipMemcpyHostToDevice));
  if (y) {
    PM_HIP(hipMalloc(&dy, 8 * n));
    PM_HIP(hipMemcpy(dy, y, 8 * n, hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(math_eval_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, 0, fn, n, dx, dy,
                     dout);
  PM_HIP(hipGetLastError());
  PM_HIP(hipMemcpy(out, dout, 8 * n, hipMemcpyDeviceToHost));
  (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  (void)hipFree(dout);
  return PELEMP_OK;
}

int pelemp_device_count(int32_t* n) {
  if (!n) return PELEMP_E_ARG;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return PELEMP_OK;
}

int pelemp_device_malloc(void** ptr, int64_t bytes) {
  if (!ptr || bytes < 0) return PELEMP_E_ARG;
  if (hipMalloc(ptr, static_cast<size_t>(bytes > 0 ? bytes : 8)) != hipSuccess) {
    set_last_error("hipMalloc failed");
    return PELEMP_E_NOMEM;
  }
  return PELEMP_OK;
}

int pelemp_device_free(void* ptr) {
  if (ptr) PM_HIP(hipFree(ptr));
  return PELEMP_OK;
}

int pelemp_memcpy_htod(void* dst, const void* src, int64_t bytes) {
  if (bytes > 0) PM_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
  return PELEMP_OK;
}

int pelemp_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
  if (bytes > 0) PM_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
  return PELEMP_OK;
}

int pelemp_memcpy_dtod(void* dst, const void* src, int64_t bytes) {
  if (bytes > 0) PM_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice));
  return PELEMP_OK;
}

int pelemp_memset_device(void* ptr, int32_t value, int64_t bytes) {
  if (bytes > 0) PM_HIP(hipMemset(ptr, value, static_cast<size_t>(bytes)));
  return PELEMP_OK;
}

int pelemp_fab_set_zero(double* ptr, int64_t n_words, void* stream) {
  if (n_words < 0 || (n_words > 0 && !ptr)) {
    set_last_error("invalid argument");
    return PELEMP_E_ARG;
  }
  if (n_words == 0) return PELEMP_OK;
  const unsigned grid = static_cast<unsigned>(std::min<long long>((n_words + 255) / 256, 16384));
  hipLaunchKernelGGL(zero_fill_kernel, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), ptr,
                     static_cast<long long>(n_words));
  PM_HIP(hipGetLastError());
  return PELEMP_OK;
}

int pelemp_device_synchronize(void) {
  PM_HIP(hipDeviceSynchronize());
  return PELEMP_OK;
}

int pelemp_stream_create(void** stream) {
  if (!stream) return PELEMP_E_ARG;
  hipStream_t s = nullptr;
  PM_HIP(hipStreamCreate(&s));  // default priority
  *stream = s;
  return PELEMP_OK;
}

int pelemp_stream_destroy(void* stream) {
  if (stream) PM_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return PELEMP_OK;
}

int pelemp_stream_synchronize(void* stream) {
  PM_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return PELEMP_OK;
}

struct pelemp_timer {
  hipEvent_t e[2];
};

int pelemp_timer_create(void** timer) {
  if (!timer) return PELEMP_E_ARG;
  auto* t = new pelemp_timer();
  if (hipEventCreate(&t->e[0]) != hipSuccess || hipEventCreate(&t->e[1]) != hipSuccess) {
    delete t;
    set_last_error("hipEventCreate failed");
    return PELEMP_E_HIP;
  }
  *timer = t;
  return PELEMP_OK;
}

int pelemp_timer_record(void* timer, int32_t which, void* stream) {
  if (!timer || which < 0 || which > 1) return PELEMP_E_ARG;
  PM_HIP(hipEventRecord(static_cast<pelemp_timer*>(timer)->e[which], static_cast<hipStream_t>(stream)));
  return PELEMP_OK;
}

int pelemp_timer_elapsed(void* timer, double* ms) {
  if (!timer || !ms) return PELEMP_E_ARG;
  auto* t = static_cast<pelemp_timer*>(timer);
  PM_HIP(hipEventSynchronize(t->e[1]));
  float f = 0.f;
  PM_HIP(hipEventElapsedTime(&f, t->e[0], t->e[1]));
  *ms = f;
  return PELEMP_OK;
}

int pelemp_timer_destroy(void* timer) {
  if (!timer) return PELEMP_OK;
  auto* t = static_cast<pelemp_timer*>(timer);
  (void)hipEventDestroy(t->e[0]);
  (void)hipEventDestroy(t->e[1]);
  delete t;
  return PELEMP_OK;
}

}  // extern "C"
