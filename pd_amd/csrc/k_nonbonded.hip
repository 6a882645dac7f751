// k_nonbonded.hip -- FF_NonBonded::calcForces on the device.
//
// Restates FF_NonBonded_CalcForces_T_fast2 (nonbonded.cpp:559-910) over the
// cluster list built by k_list.hip.  One wavefront owns one i-cluster (8 atoms,
// held in scalar registers); its 64 lanes walk the cluster's list entries, one
// j atom per lane, and loop over the 8 i atoms.  Per pair:
//   * displacement and the cutoff test exactly as PD in FP64:
//       d = (p_j - p_i) + basis[image],  r^2 = (x*x + y*y) + z*z,  keep r^2 <= Cutoff^2
//     so the interacting pair set of every step is PD's, bit for bit;
//   * the pair energy/force in FP32 (kcal/mol, A) with PD's expressions:
//     LJ 2eps(A/2 - B) with its energy switch (always on, nonbonded.cpp:748-761),
//     Coulomb plain / energy-switched / force-switched (nonbonded.cpp:766-821);
//   * 1-2/1-3 pairs (scale 0) drop out, 1-4 pairs use Vdw14/Elec14Scaling
//     (nonbonded.cpp:650-651).
// The j force of a lane is complete after the 8 i atoms and is added once, in
// 64-bit fixed point; i forces stay in registers until the end of the cluster.
// Fixed-point integer adds are order-independent, so the forces are bitwise
// reproducible.  Energies go to per-wave FP64 partials reduced in a fixed order.
#include "engine.hpp"

#include <cmath>

namespace pdgpu {

namespace {

enum { EL_NONE = 0, EL_PLAIN = 1, EL_ESWITCH = 3, EL_FSWITCH = 4 };

__global__ void k_gather_slots(int nslots, const int *__restrict__ n_clusters, const int *__restrict__ slot_orig,
                               const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, double *__restrict__ sx, double *__restrict__ sy,
                               double *__restrict__ sz) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots || s >= *n_clusters * kClusterSize) return;
  int a = slot_orig[s];
  if (a >= 0) {
    sx[s] = x[a];
    sy[s] = y[a];
    sz[s] = z[a];
  } else {  // padding slot: far away, zero parameters, never listed
    sx[s] = 1.0e6;
    sy[s] = 1.0e6;
    sz[s] = 1.0e6;
  }
}

__global__ void k_gather_params(int nslots, const int *__restrict__ n_clusters, const int *__restrict__ slot_orig,
                                const float *__restrict__ q, const float *__restrict__ r,
                                const float *__restrict__ e, float *__restrict__ sq, float *__restrict__ sr,
                                float *__restrict__ se) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots || s >= *n_clusters * kClusterSize) return;
  int a = slot_orig[s];
  sq[s] = a >= 0 ? q[a] : 0.0f;
  sr[s] = a >= 0 ? r[a] : 0.0f;
  se[s] = a >= 0 ? e[a] : 0.0f;
}

__global__ void k_zero_ll(long long *__restrict__ a, long long *__restrict__ b, long long *__restrict__ c,
                          int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    a[i] = 0;
    b[i] = 0;
    c[i] = 0;
  }
}

// Image index (i+1)*9 + (j+1)*3 + (k+1) -> the basis vector (i*Lx, j*Ly, k*Lz)
// of PeriodicBox::setupCellDimensions (space.cpp:94-109).
__device__ __forceinline__ double image_shift(int img, int axis, double L) {
  int k = axis == 0 ? img / 9 : (axis == 1 ? (img / 3) % 3 : img % 3);
  return k == 1 ? 0.0 : (k == 2 ? L : -L);
}

template <int ELEC, bool VDW>
__global__ __launch_bounds__(64) void k_nonbonded(
    NBKernelParams P, const int *__restrict__ n_clusters, int grid_clusters, double Lx, double Ly,
    double Lz, const double *__restrict__ sx, const double *__restrict__ sy, const double *__restrict__ sz,
    const float *__restrict__ sq, const float *__restrict__ sr, const float *__restrict__ se,
    const uint2 *__restrict__ entries, int stride, const int *__restrict__ cl_count,
    long long *__restrict__ fx, long long *__restrict__ fy, long long *__restrict__ fz,
    double *__restrict__ e_partial) {
  // XCD-aware order: consecutive workgroups are spread round-robin over the 8
  // XCDs, so give each XCD a contiguous run of (spatially adjacent) clusters.
  const int nblk = gridDim.x;
  const int per = nblk / 8;
  const int ci = (per * 8 == nblk) ? (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  const int lane = threadIdx.x;
  const int ncl = *n_clusters;
  if (ci >= ncl) {
    if (lane < 4) e_partial[4 * ci + lane] = 0.0;
    return;
  }
  double xi[8], yi[8], zi[8];
  float qi[8], ri[8], ei[8];
#pragma unroll
  for (int ii = 0; ii < 8; ii++) {
    const int s = ci * kClusterSize + ii;
    xi[ii] = sx[s];
    yi[ii] = sy[s];
    zi[ii] = sz[s];
    qi[ii] = sq[s];
    ri[ii] = sr[s];
    ei[ii] = se[s];
  }
  float fix[8], fiy[8], fiz[8];
#pragma unroll
  for (int ii = 0; ii < 8; ii++) fix[ii] = fiy[ii] = fiz[ii] = 0.0f;
  double ev = 0.0, ee = 0.0, vir = 0.0;  // per-lane FP64 energy accumulators

  const int cnt = cl_count[ci];
  const uint2 *E = entries + (size_t)ci * stride;
  for (int k0 = 0; k0 < cnt; k0 += 64) {
    const int k = k0 + lane;
    const bool valid = k < cnt;
    const uint2 ent = valid ? E[k] : make_uint2(0u, 0u);
    const int j = (int)(ent.x & 0xFFFFFFu);
    const int img = (int)(ent.x >> 24);
    const unsigned lm = ent.y & 0xFFu;
    const unsigned codes = ent.y >> 8;
    const unsigned clo = codes & 0x5555u, chi = (codes >> 1) & 0x5555u;
    const unsigned excl = clo ^ chi;  // codes 1 and 2: 1-2 / 1-3 pairs, scale 0
    const unsigned is14 = clo & chi;  // code 3: 1-4 pair
    const double xj = sx[j], yj = sy[j], zj = sz[j];
    const double shx = image_shift(img, 0, Lx), shy = image_shift(img, 1, Ly), shz = image_shift(img, 2, Lz);
    const float qj = sq[j], rj = sr[j], ej = se[j];
    float fjx = 0.0f, fjy = 0.0f, fjz = 0.0f;
    float ev8 = 0.0f, ee8 = 0.0f, vir8 = 0.0f;  // the 8 pairs of this entry
#pragma unroll
    for (int ii = 0; ii < 8; ii++) {
      double dx, dy, dz, r2;
      {
#pragma clang fp contract(off)
        dx = (xj - xi[ii]) + shx;
        dy = (yj - yi[ii]) + shy;
        dz = (zj - zi[ii]) + shz;
        r2 = dx * dx + dy * dy + dz * dz;
      }
      const bool on = ((lm >> ii) & 1u) && !((excl >> (2 * ii)) & 1u) && (r2 <= P.cut2);
      const bool s14 = (is14 >> (2 * ii)) & 1u;
      const float r2f = on ? (float)r2 : 1.0f;
      const float rinv = __builtin_amdgcn_rsqf(r2f);
      const float r = r2f * rinv;
      const float rinv2 = rinv * rinv;
      float dEdr = 0.0f, Ev = 0.0f, Ee = 0.0f;
      if (VDW) {
        const float eps = (s14 ? P.vdw14 : 1.0f) * ei[ii] * ej;
        const float Rij = ri[ii] + rj;
        const float s2 = Rij * Rij * rinv2;
        const float B = s2 * s2 * s2;
        const float A = B * B;
        float e0 = eps * (A - 2.0f * B);
        float d0 = -12.0f * eps * rinv * (A - B);
        const float t = fmaxf(r - P.vdwinner, 0.0f);
        const float S0 = 1.0f - (P.invw_vdw * t) * (P.invw_vdw * t);
        const float dS = -P.invw2_vdw4 * t * S0;
        const float S = S0 * S0;
        d0 = S * d0 + e0 * dS;
        e0 = e0 * S;
        const bool vin = r2 < P.vdwcut2;
        Ev = vin ? e0 : 0.0f;
        dEdr += vin ? d0 : 0.0f;
      }
      if (ELEC != EL_NONE) {
        const float qq = (s14 ? P.elec14 : 1.0f) * qi[ii] * qj;
        if (ELEC == EL_PLAIN) {
          Ee = qq * rinv;
          dEdr += -Ee * rinv;
        } else if (ELEC == EL_ESWITCH) {
          float e0 = qq * rinv;
          float d0 = -e0 * rinv;
          const float t = fmaxf(r - P.inner, 0.0f);
          const float S0 = 1.0f - (P.invw_el * t) * (P.invw_el * t);
          const float dS = -P.invw2_el4 * t * S0;
          const float S = S0 * S0;
          d0 = S * d0 + e0 * dS;
          Ee = e0 * S;
          dEdr += d0;
        } else {  // force switch (Steinbach & Brooks), nonbonded.cpp:798-819
          const bool outer = r > P.inner;
          const float dc = P.rc2 - r2f;
          const float d_out = -qq * P.sA * rinv2 * dc * dc * (P.rc2 - 3.0f * P.rin2 + 2.0f * r2f);
          const float e_out = -qq * (P.sA * r * (P.sB * rinv2 + P.sC + P.sD * r2f + P.sE * r2f * r2f) - P.fcut);
          const float d_in = -qq * rinv2;
          const float e_in = qq * (rinv - P.eshift);
          Ee = outer ? e_out : e_in;
          dEdr += outer ? d_out : d_in;
        }
      }
      const float sc = on ? dEdr * rinv : 0.0f;
      const float gx = (float)dx * sc, gy = (float)dy * sc, gz = (float)dz * sc;
      fix[ii] += gx;
      fiy[ii] += gy;
      fiz[ii] += gz;
      fjx -= gx;
      fjy -= gy;
      fjz -= gz;
      ev8 += on ? Ev : 0.0f;
      ee8 += on ? Ee : 0.0f;
      vir8 += on ? r * dEdr : 0.0f;
    }
    ev += (double)ev8;
    ee += (double)ee8;
    vir += (double)vir8;
    if (valid && lm) {
      fix_add(&fx[j], fix_from_float(fjx));
      fix_add(&fy[j], fix_from_float(fjy));
      fix_add(&fz[j], fix_from_float(fjz));
    }
  }
#pragma unroll
  for (int ii = 0; ii < 8; ii++) {
    const float gx = wave_sum_f(fix[ii]);
    const float gy = wave_sum_f(fiy[ii]);
    const float gz = wave_sum_f(fiz[ii]);
    if (lane == ii) {
      const int s = ci * kClusterSize + ii;
      fix_add(&fx[s], fix_from_float(gx));
      fix_add(&fy[s], fix_from_float(gy));
      fix_add(&fz[s], fix_from_float(gz));
    }
  }
  const double sev = wave_sum(ev);
  const double see = wave_sum(ee);
  const double svir = wave_sum(vir);
  if (lane == 0) {
    e_partial[4 * ci + 0] = sev;
    e_partial[4 * ci + 1] = see;
    e_partial[4 * ci + 2] = svir;
    e_partial[4 * ci + 3] = 0.0;
  }
}

template <int ELEC, bool VDW>
void launch(Engine &E, int grid) {
  hipLaunchKernelGGL((k_nonbonded<ELEC, VDW>), dim3(grid), dim3(64), 0, E.stream, E.nbk, E.n_clusters_d.p,
                     grid, E.periodic ? E.box[0] : 0.0, E.periodic ? E.box[1] : 0.0,
                     E.periodic ? E.box[2] : 0.0, E.sx.p, E.sy.p, E.sz.p, E.sq.p, E.sr.p, E.se.p, E.entries.p,
                     (int)E.entry_capacity, E.cl_entry_count.p, E.nbx.p, E.nby.p, E.nbz.p, E.e_partial.p);
}

}  // namespace

void reduce4(hipStream_t s, const double *part, int count, double *out);

void Engine::setup_nonbonded(const pdgpu_nb_params &p) {
  require(!h_charge.empty(), "atom parameters not set");
  // FF_NonBonded::setupBasisVectors refuses more than 32 image vectors (nonbonded.cpp:1059-1069);
  // PeriodicBox::setupCellDimensions(Cutoff) gives (2n+1)^3 vectors, n = ceil(Cutoff/half box).
  const double c = std::max(p.Cutoff, p.VdwCutoff);
  if (periodic) {
    for (int d = 0; d < 3; d++) {
      double nx = std::ceil(std::max(list_cutoff, c) / (box[d] / 2.0));
      if (nx > 1.0)
        throw EngineError(PDGPU_ERR_SETUP,
                          "FF_NonBonded: Cannot use more than 32 periodic boundary image vectors "
                          "(cutoff greater than half the periodic box width)");
    }
  }
  nb = p;
  list_cutoff = std::max(list_cutoff, c);  // NeighbourListBase::requestCutoff: the maximum wins
  NBKernelParams k{};
  const double C = p.Cutoff, I = p.InnerCutoff;
  k.cut2 = C * C;
  k.vdwcut2 = p.VdwCutoff * p.VdwCutoff;
  k.cut = (float)C;
  k.inner = (float)I;
  k.vdwcut = (float)p.VdwCutoff;
  k.vdwinner = (float)p.VdwInnerCutoff;
  // A zero switching width never switches (no pair has Inner < r <= Cutoff); keep
  // the device constants finite so the branch-free switch stays exactly 1.
  const double vw = p.VdwCutoff - p.VdwInnerCutoff;
  k.invw_vdw = vw > 0 ? (float)(1.0 / vw) : 0.0f;
  k.invw2_vdw4 = vw > 0 ? (float)(4.0 / (vw * vw)) : 0.0f;
  const double ew = C - I;
  k.invw_el = ew > 0 ? (float)(1.0 / ew) : 0.0f;
  k.invw2_el4 = ew > 0 ? (float)(4.0 / (ew * ew)) : 0.0f;
  // force-switch constants, nonbonded.cpp:620-636
  auto sqr = [](double v) { return v * v; };
  auto cube = [](double v) { return v * v * v; };
  const double sA = 1.0 / cube(sqr(C) - sqr(I));
  const double sB = -(cube(sqr(C)) - 3.0 * sqr(C) * sqr(C) * sqr(I));
  const double sC = 6.0 * sqr(C) * sqr(I);
  const double sD = -(sqr(C) + sqr(I));
  const double sE = 2.0 / 5.0;
  const double finner = sA * (sB * (1.0 / I) + sC * I + sD * cube(I) + sE * cube(I) * sqr(I));
  const double fcut = sA * (sB * (1.0 / C) + sC * C + sD * cube(C) + sE * cube(C) * sqr(C));
  const double eshift = (I < C) ? 1 / I + (finner - fcut) : 0.0;
  k.sA = (float)sA;
  k.sB = (float)sB;
  k.sC = (float)sC;
  k.sD = (float)sD;
  k.sE = (float)sE;
  k.fcut = (float)fcut;
  k.eshift = (float)eshift;
  k.rc2 = (float)(C * C);
  k.rin2 = (float)(I * I);
  k.vdw14 = (float)p.Vdw14Scaling;
  k.elec14 = (float)p.Elec14Scaling;
  k.vdw_on = p.DoVdw ? 1 : 0;
  int mode = 1;  // dispatch as FF_NonBonded::calcForces, nonbonded.cpp:1228-1238
  if (p.EnergySwitch) mode = 3;
  if (p.ForceSwitch) mode = 4;
  if (!p.DoElec) mode = 0;
  k.elec_mode = mode;
  if (p.VdwCor) {
    const double R = p.VdwCorRadius;
    k.vdwcor_per_atom = -0.5 * p.VdwCorDensity * p.VdwCorEpsilon * (8.0 / 3.0) * Consts::PI * sqr(R) *
                        sqr(R) * sqr(R) / cube(p.VdwCorCutoff);  // kcal/mol
  }
  nbk = k;
  // per-atom parameters in kernel units: charge * sqrt(econv/Dielectric), sqrt(eps[kcal/mol])
  std::vector<float> q(n), r(n), e(n);
  const double qs = std::sqrt(Consts::econv() / p.Dielectric);
  const double es = 1.0 / std::sqrt(Consts::kcal_to_J());
  for (int i = 0; i < n; i++) {
    q[i] = (float)(h_charge[i] * qs);
    r[i] = (float)h_radius[i];
    e[i] = (float)(h_epsilon[i] * es);
  }
  q_scaled.upload(q, stream);
  lj_r.upload(r, stream);
  lj_e.upload(e, stream);
  have_nb = true;
  list_valid = false;
}

void Engine::gather_positions() {
  const int nslots = max_clusters * kClusterSize;
  sx.resize(nslots);
  sy.resize(nslots);
  sz.resize(nslots);
  hipLaunchKernelGGL(k_gather_slots, dim3((nslots + 255) / 256), dim3(256), 0, stream, nslots, n_clusters_d.p,
                     slot_orig.p, x.p, y.p, z.p, sx.p, sy.p, sz.p);
}

// Per-slot parameters and accumulators; called once per list rebuild.
void Engine::prepare_slots() {
  const int nslots = max_clusters * kClusterSize;
  sq.resize(nslots);
  sr.resize(nslots);
  se.resize(nslots);
  nbx.resize(nslots);
  nby.resize(nslots);
  nbz.resize(nslots);
  if (have_nb)
    hipLaunchKernelGGL(k_gather_params, dim3((nslots + 255) / 256), dim3(256), 0, stream, nslots,
                       n_clusters_d.p, slot_orig.p, q_scaled.p, lj_r.p, lj_e.p, sq.p, sr.p, se.p);
}

void Engine::launch_nonbonded(double *energy_out) {
  require(have_nb, "nonbonded component not set up");
  require(list_valid, "neighbour list not built");
  const int nslots = max_clusters * kClusterSize;
  hipLaunchKernelGGL(k_zero_ll, dim3((nslots + 255) / 256), dim3(256), 0, stream, nbx.p, nby.p, nbz.p, nslots);
  gather_positions();
  const int grid = (max_clusters + 7) / 8 * 8;
  e_partial.resize((size_t)grid * 4);
  const bool v = nbk.vdw_on != 0;
  {
    ScopedTimer t(*this, "nonbonded");
    switch (nbk.elec_mode) {
      case 0: v ? launch<EL_NONE, true>(*this, grid) : launch<EL_NONE, false>(*this, grid); break;
      case 1: v ? launch<EL_PLAIN, true>(*this, grid) : launch<EL_PLAIN, false>(*this, grid); break;
      case 3: v ? launch<EL_ESWITCH, true>(*this, grid) : launch<EL_ESWITCH, false>(*this, grid); break;
      default: v ? launch<EL_FSWITCH, true>(*this, grid) : launch<EL_FSWITCH, false>(*this, grid); break;
    }
  }
  PDGPU_HIP(hipGetLastError());
  reduce4(stream, e_partial.p, grid, energy_out);
}

}  // namespace pdgpu
