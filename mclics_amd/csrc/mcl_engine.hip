// libmclgpu: checkerboard Metropolis sweep, FP64 observables and their C ABI, written for MI355X (gfx950).
//
// Device layout (per lattice, pooled over the lattices of a context):
//   The lattice is split into its two checkerboard colours c = (i + j + k) & 1. Site (i, j, k) of colour c lives at
//   position t = (i >> 1) + nxh * (j + ny * k) of that colour's block (nxh = ceil(nx / 2)), in three FP32 arrays
//   X, Y, Z (structure of arrays). All neighbours of a site have the other colour, and the six neighbours of the
//   sites t, t+1, ... of one row are consecutive entries of rows of the other colour, so every load of a sweep
//   pass is a coalesced 4-byte-per-lane access. Point types (uint8) and the per-site anchoring vector follow the
//   same layout and exist only when a lattice has empty or surface sites.
//
// One sweep = two kernel launches (colour 0, then colour 1). Each thread makes one Metropolis trial on one site of
// the active colour (proposal of evolve.cu:162-182, energies of geometry.cu:25-120 / potential.cu:11-58, rule of
// evolve.cu:212-239). Random numbers come from Philox4x32-10 keyed by (seed, lattice) with counter
// (site, sweep), so the generator has no state in memory. Accepted rotations are counted with a wave ballot and
// one atomic per workgroup into 64 striped counters per lattice (a ring of three sweeps); the colour-0 pass of the
// next sweep sums the 64 counters in every wave and applies the ang_var rule of evolve.cu:430-436 on the device,
// so the rule is exact and needs neither a host round trip nor an extra launch per sweep.
//
// A measurement is one pass over both colours that accumulates, in FP64, the site energy (get_Energy,
// evolve.cu:66-108), the Q-tensor sums (parameter_order.cpp:40-71), the radial order (parameter_order.cpp:223-245)
// and the pear observables (parameter_order.cpp:131-215) into one partial per workgroup; a finalize kernel sums the
// partials of each lattice in a fixed order (bitwise reproducible), solves the 3x3 eigenproblem for S in FP64 and
// adds the sample to the lattice's FP64 moments.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mclgpu.h"
#include "philox.h"

namespace {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kFinBlock = 256;  // threads of the finalize kernel
constexpr int kFinWaves = kFinBlock / 64;
constexpr int kMeasPerThread = 8;    // sites per thread in a measurement workgroup
constexpr int kNQ = 24;              // FP64 quantities per measurement partial
constexpr uint32_t kTagSweep = 0x4D435357u;  // counter word 3 of the sweep draws

// Lattice flags.
constexpr int F_PX = 1, F_PY = 2, F_PZ = 4, F_MASK = 8, F_SURF = 16, F_RSPH = 32, F_RCYL = 64, F_SURF1 = 128;

// Division by a runtime constant d for 0 <= n < 2^31: q = (n * m) >> sh with l = ceil(log2 d),
// m = ceil(2^(31+l) / d), sh = 31 + l (Granlund & Montgomery 1994; Hacker's Delight 10-8: exact because
// m d - 2^(31+l) < d <= 2^l).
struct FastDiv {
  uint32_t m, sh;
};
FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  uint64_t m = ((1ull << (31 + l)) + d - 1) / d;
  return FastDiv{(uint32_t)m, 31 + l};
}
__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t m, uint32_t sh) {
  return (uint32_t)(((uint64_t)n * m) >> sh);
}

// One anchoring (surface) type. Device anchoring kernels (the GPU selection of evolve.cu:280-297): kind 0 = -W (n.v)^2
// (homeotropic, rp, rp_tangent, with v the per-site vector), 1 = +W (n.v)^2 (fg), 2 = W U_GHRL(n, v, v),
// 3 = W U_GHRL(n, axis, v), 4 = W U_GHRL(n, tangent part of n, v).
struct SurfDev {
  int kind;
  float W, ax, ay, az;
};

struct LatDev {
  int nx, ny, nz, nxh, nc;  // nc = sites per colour block = nxh * ny * nz
  uint32_t mxh, sxh, my, sy;
  uint32_t plane, ywrap, zwrap;  // nxh * ny, (ny - 1) * nxh, (nz - 1) * plane: colour-block strides
  int flags;
  int nvalid;
  int kc;         // acceptance counters in use (a power of two up to kCnt; set at context creation)
  uint32_t off;   // pooled offset of the colour-0 block; colour 1 follows at off + nc (pool < 2^31 sites)
  uint32_t key0, key1;
  int surf0;      // first entry of this lattice in the pooled surface table
  int mp0, mp1;   // measurement partials [mp0, mp1)
  float A, B1, C, gl, gm, gn, gr, gs;
  float ex, ey, ez, elecA, elecP;
  float cx, cy, cz;  // centre of the radial observables (Nx/2, Ny/2, Nz/2 with integer division)
  // F_SURF1: every surface of the lattice is this one (the common case: one anchoring per lattice), so the trial
  // takes it from the lattice's (wave-uniform) record instead of loading a table entry per lane and switching on a
  // per-lane kind. Sphere lattices 14 % faster per pass, the GHRL slab 3.5 %.
  SurfDev s1;
};

// Per-lattice control block. av_store[p] is the ang_var used by the sweeps of parity p (written by the lattice's
// first colour-0 workgroup); a value set by the host (av_set) replaces the rule at sweep av_set_sweep.
struct CtrlDev {
  float T, invT, elecE, av_set;
  float av_store[2];
  unsigned long long av_set_sweep;
};

// Striped acceptance counters: kCnt per lattice and sweep, each in its own 128-byte cache line (same-address or
// same-line device-scope atomics from thousands of workgroups serialize at about 12 ns each).
constexpr int kCnt = 64;
constexpr int kCntStride = 32;                     // uint32 per counter slot (128 bytes)
constexpr int kCntSweep = kCnt * kCntStride;       // one sweep's counters
constexpr int kCntLattice = 3 * kCntSweep;         // ring of three sweeps

struct MomDev {
  double n, S, S2, E, E2, srs, prs, src, prc, P, dth, divn;
};
struct ObsDev {
  double nvalid, E, Q[6], S, dir[3], srs, prs, src, prc, P, dth, divn;
};

struct V3 {
  float x, y, z;
};
// The FP32 expressions of the trial are written with explicit fmaf wherever two products are summed, so that
// every kernel instance rounds them the same way (the compiler's choice of which product to contract could
// otherwise differ between instances, and trajectories must not depend on the instance).
__device__ __forceinline__ float dot3(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
template <int AX>
__device__ __forceinline__ float comp(V3 a) {
  return AX == 0 ? a.x : (AX == 1 ? a.y : a.z);
}

// Loads and stores at a 32-bit element index of a uniform base pointer. The pool of a context holds fewer than
// 2^30 sites, so the byte offset fits in 32 bits and the compiler can use the scalar-base + 32-bit-offset form.
template <typename T>
__device__ __forceinline__ T ldx(const T *__restrict__ base, uint32_t idx) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (idx * (uint32_t)sizeof(T)));
}
// The sweep's director stores are non-temporal (streaming hint: the pass never reads what it writes, the next pass
// reads it after the kernel boundary); they also take the scalar-base form. 128^3: +3 % over plain stores.
template <typename T>
__device__ __forceinline__ void stx(T *__restrict__ base, uint32_t idx, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (idx * (uint32_t)sizeof(T))));
}

// ---- pair potentials (potential.cpp:19-56 / potential.cu:11-43) ----

// GHRL with a general bond vector r (used by the GHRL anchorings, geometry.cu:377-390).
__device__ __forceinline__ float ghrl_general(const LatDev &L, V3 ni, V3 nj, V3 r) {
  const float ai = dot3(ni, r), aj = dot3(nj, r), nij = dot3(ni, nj);
  const float pij = 1.5f * nij * nij - 0.5f;
  const float cross =
      (ni.z * nj.y - ni.y * nj.z) * r.x + (ni.x * nj.z - ni.z * nj.x) * r.y + (ni.y * nj.x - ni.x * nj.y) * r.z;
  const float E1 = 1.5f * ai * ai + 1.5f * aj * aj - 1.0f;
  return E1 * (L.gr * pij + L.gl) + L.gm * (ai * aj * nij - (1.0f / 9.0f)) + L.gn * pij +
         L.gs * (nij > 0.0f ? cross : -cross);
}

// Bond energy with the neighbour in direction D (0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z); r = +-e_axis is a
// compile-time constant, so a_i, a_j and the cross product reduce to single components. For LL the caller
// multiplies the sum of nij^2 by -1.5 A.
template <int POT, int D>
__device__ __forceinline__ float bond(const LatDev &L, V3 a, V3 b) {
  constexpr int AX = D >> 1;
  constexpr float SG = (D & 1) ? -1.0f : 1.0f;
  const float nij = dot3(a, b);
  if constexpr (POT == MCL_POT_LL) {
    return nij * nij;
  } else if constexpr (POT == MCL_POT_GHRL) {
    // the achiral part; a_i = r.n_i and a_j = r.n_j enter only as a_i^2, a_j^2 and a_i a_j, so the bond sign drops
    // out. The chiral term sigma (+-c) is bond_chiral, summed separately by bond_sum.
    const float ai = comp<AX>(a), aj = comp<AX>(b);
    const float pij = 1.5f * nij * nij - 0.5f;
    const float E1 = 1.5f * fmaf(aj, aj, ai * ai) - 1.0f;
    const float e = fmaf(E1, fmaf(L.gr, pij, L.gl), L.gm * fmaf(ai * aj, nij, -(1.0f / 9.0f)));
    return fmaf(L.gn, pij, e);
  } else {
    // pear: -A b^2 - B1 b - (C/4) (1 + b)^2 r.(n_j - n_i)
    const float dr = SG * (comp<AX>(b) - comp<AX>(a));
    const float op = 1.0f + nij;
    return fmaf(-(0.25f * L.C) * (op * op), dr, fmaf(-L.A * nij, nij, -(L.B1 * nij)));
  }
}

__device__ __forceinline__ float surface_energy(const LatDev &L, const SurfDev &s, V3 n, V3 v) {
  switch (s.kind) {
    case 0: {
      const float d = dot3(n, v);
      return -s.W * d * d;
    }
    case 1: {
      const float d = dot3(n, v);
      return s.W * d * d;
    }
    case 2:
      return s.W * ghrl_general(L, n, v, v);
    case 3:
      return s.W * ghrl_general(L, n, V3{s.ax, s.ay, s.az}, v);
    default: {
      const float d = dot3(n, v);
      V3 t{n.x - d * v.x, n.y - d * v.y, n.z - d * v.z};
      const float mod = sqrtf(dot3(t, t));
      if (mod < 1e-6f) return 0.0f;
      t.x /= mod;
      t.y /= mod;
      t.z /= mod;
      return s.W * ghrl_general(L, n, t, v);
    }
  }
}

// Local environment of one site (the nni fullni[8] of define.h:24-33, in registers).
struct Site {
  V3 n;
  V3 nb[6];
  V3 sv;
  unsigned ex;  // bit d: neighbour d exists and is not empty
  int ptv;
  int i, j, k;
  uint32_t qx[2];  // NBX: pool indices of the +x and -x neighbours
};

// GHRL chiral factor of bond D: +-c with c = r.(n_j x n_i) (potential.cpp:50-52), the sign of n_i.n_j.
template <int D>
__device__ __forceinline__ float bond_chiral(V3 a, V3 b) {
  constexpr int AX = D >> 1;
  constexpr float SG = (D & 1) ? -1.0f : 1.0f;
  const float cr = SG * (AX == 0 ? fmaf(a.z, b.y, -(a.y * b.z))
                                 : (AX == 1 ? fmaf(a.x, b.z, -(a.z * b.x)) : fmaf(a.y, b.x, -(a.x * b.y))));
  return dot3(a, b) > 0.0f ? cr : -cr;
}

// Bond part of the site energy: the bonds to the existing neighbours (LL scaled by -1.5 A).
template <int POT>
__device__ __forceinline__ float bond_sum(const LatDev &L, const Site &s, V3 n) {
  float e = 0.0f;
  if constexpr (POT == MCL_POT_LL) {
    // sum of nij^2 as a chain of fmaf (the same rounding whether or not the existence mask is a constant)
#pragma unroll
    for (int d = 0; d < 6; d++) {
      const float q = dot3(n, s.nb[d]);
      e = (s.ex & (1u << d)) ? fmaf(q, q, e) : e;
    }
    return e * (-1.5f * L.A);
  }
  // Every bond is evaluated and then summed in direction order: a plain sum when all six neighbours exist (the
  // bulk of any lattice), and a sum of selected terms otherwise, where a missing neighbour adds +0.0f. Both add the
  // same values in the same order. (With the bond expressions inside the conditionals the compiler branches around
  // each one under an exec mask, and per-bond selects on every site cost a quarter of the bond arithmetic.)
  const float b0 = bond<POT, 0>(L, n, s.nb[0]), b1 = bond<POT, 1>(L, n, s.nb[1]);
  const float b2 = bond<POT, 2>(L, n, s.nb[2]), b3 = bond<POT, 3>(L, n, s.nb[3]);
  const float b4 = bond<POT, 4>(L, n, s.nb[4]), b5 = bond<POT, 5>(L, n, s.nb[5]);
  if (s.ex == 63u) {
    e = ((((b0 + b1) + b2) + b3) + b4) + b5;
  } else {
    e += (s.ex & 1u) ? b0 : 0.0f;
    e += (s.ex & 2u) ? b1 : 0.0f;
    e += (s.ex & 4u) ? b2 : 0.0f;
    e += (s.ex & 8u) ? b3 : 0.0f;
    e += (s.ex & 16u) ? b4 : 0.0f;
    e += (s.ex & 32u) ? b5 : 0.0f;
  }
  if constexpr (POT == MCL_POT_GHRL) {
    // sigma = 0 without chirality (setGHRL, io.cpp:340-360): a uniform branch skips the cross products. With
    // sigma != 0 the chiral factors are summed and added as sigma x sum (fmaf(0, c, e) would be e exactly).
    if (L.gs != 0.0f) {
      const float c0 = bond_chiral<0>(n, s.nb[0]), c1 = bond_chiral<1>(n, s.nb[1]);
      const float c2 = bond_chiral<2>(n, s.nb[2]), c3 = bond_chiral<3>(n, s.nb[3]);
      const float c4 = bond_chiral<4>(n, s.nb[4]), c5 = bond_chiral<5>(n, s.nb[5]);
      float c = 0.0f;
      if (s.ex == 63u) {
        c = ((((c0 + c1) + c2) + c3) + c4) + c5;
      } else {
        c += (s.ex & 1u) ? c0 : 0.0f;
        c += (s.ex & 2u) ? c1 : 0.0f;
        c += (s.ex & 4u) ? c2 : 0.0f;
        c += (s.ex & 8u) ? c3 : 0.0f;
        c += (s.ex & 16u) ? c4 : 0.0f;
        c += (s.ex & 32u) ? c5 : 0.0f;
      }
      e = fmaf(L.gs, c, e);
    }
  }
  return e;
}

// One-site part of the site energy: anchoring on surface sites and the field.
template <bool SURF, bool FIELD>
__device__ __forceinline__ float site_terms(const LatDev &L, int ptv, V3 sv, V3 n, const SurfDev *__restrict__ surf,
                                            float elecE) {
  float e = 0.0f;
  if constexpr (SURF) {
    if (ptv > 1) e += (L.flags & F_SURF1) ? surface_energy(L, L.s1, n, sv)
                                          : surface_energy(L, surf[L.surf0 + ptv - 2], n, sv);
  }
  if constexpr (FIELD) {
    const float nE = elecE * fmaf(n.z, L.ez, fmaf(n.y, L.ey, n.x * L.ex));
    e += fmaf(-L.elecA * nE, nE, -(L.elecP * nE));
  }
  return e;
}

// Site energy: bonds + anchoring + field (Bulk/Slab/Sphere/Custom_site_energy_GPU, geometry.cu:63-120).
template <int POT, bool SURF, bool FIELD>
__device__ __forceinline__ float site_energy(const LatDev &L, const Site &s, V3 n, const SurfDev *__restrict__ surf,
                                             float elecE) {
  float e = bond_sum<POT>(L, s, n);
  if constexpr (SURF || FIELD) e += site_terms<SURF, FIELD>(L, s.ptv, s.sv, n, surf, elecE);
  return e;
}

// Decodes colour-block position t of lattice L into (i, j, k) and loads the site and its neighbours.
// Returns false for padding positions (i >= nx) and empty sites.
// FREE = false: every axis of every lattice of the context is periodic (and MASK is false), so every neighbour
// exists and the wrap is the only boundary case.
// NBX (measuring kernels of masked lattices): also keeps the pool indices of the two x neighbours, whose one-site
// terms the measurement adds for the colour-0 sites a colour-1 site counts.
// NBPT = false (the LL sweep): the neighbours' point types are not loaded. An LL context stores its empty sites as
// zero vectors, whose bonds add fmaf(0, 0, e) = e to the LL sum, exactly what the existence mask would give.
template <bool MASK, bool SURF, bool FREE = true, bool NBX = false, bool NBPT = true>
__device__ __forceinline__ bool gather(const LatDev &L, int colour, int t, const float *__restrict__ X,
                                       const float *__restrict__ Y, const float *__restrict__ Z,
                                       const uint8_t *__restrict__ PT, const float *__restrict__ SX,
                                       const float *__restrict__ SY, const float *__restrict__ SZ, Site &s,
                                       uint32_t &self) {
  const int nx = L.nx, ny = L.ny, nz = L.nz, nxh = L.nxh;
  const uint32_t row = fdiv((uint32_t)t, L.mxh, L.sxh);
  const int m = t - (int)row * nxh;
  const int k = (int)fdiv(row, L.my, L.sy);
  const int j = (int)row - k * ny;
  if (m < 0 || m >= nxh || j < 0 || j >= ny || k >= nz) return false;  // unreachable for t < nc
  const int i = 2 * m + ((j + k + colour) & 1);
  if (i >= nx) return false;
  s.i = i;
  s.j = j;
  s.k = k;
  const uint32_t base = L.off + (uint32_t)(colour * L.nc);
  const uint32_t obase = L.off + (uint32_t)((1 - colour) * L.nc);
  self = base + (uint32_t)t;
  s.ptv = 1;
  if constexpr (MASK) {
    s.ptv = ldx(PT, self);
    if (s.ptv == 0) return false;
  }
  s.n = V3{ldx(X, self), ldx(Y, self), ldx(Z, self)};
  // Periodic_Boundary / Free_Boundary (potential.cu:50-58). A missing neighbour keeps an in-bounds index (the
  // site's own column in the other colour) and is masked out of the energy, so all loads are unconditional.
  unsigned ex = 63u;
  const uint32_t col = obase + (uint32_t)t, rowstart = obase + (uint32_t)(t - m);
  const uint32_t rs = (uint32_t)nxh, pl = L.plane;
  uint32_t nidx[6];
  if constexpr (!FREE) {
    nidx[0] = rowstart + (uint32_t)((i + 1 == nx ? 0 : i + 1) >> 1);
    nidx[1] = rowstart + (uint32_t)((i == 0 ? nx - 1 : i - 1) >> 1);
    nidx[2] = j == ny - 1 ? col - L.ywrap : col + rs;
    nidx[3] = j == 0 ? col + L.ywrap : col - rs;
    nidx[4] = k == nz - 1 ? col - L.zwrap : col + pl;
    nidx[5] = k == 0 ? col + L.zwrap : col - pl;
  } else {
    const bool px = L.flags & F_PX, py = L.flags & F_PY, pz = L.flags & F_PZ;
    int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1, km = k - 1;
    if (ip == nx) { ip = px ? 0 : i; ex &= px ? 63u : ~1u; }
    if (im < 0) { im = px ? nx - 1 : i; ex &= px ? 63u : ~2u; }
    if (jp == ny) { jp = py ? 0 : j; ex &= py ? 63u : ~4u; }
    if (jm < 0) { jm = py ? ny - 1 : j; ex &= py ? 63u : ~8u; }
    if (kp == nz) { kp = pz ? 0 : k; ex &= pz ? 63u : ~16u; }
    if (km < 0) { km = pz ? nz - 1 : k; ex &= pz ? 63u : ~32u; }
    // neighbours in y and z are whole rows / planes away in the other colour block; at a free face the index
    // stays on the site's own column (masked out of the energy)
    nidx[0] = rowstart + (uint32_t)(ip >> 1);
    nidx[1] = rowstart + (uint32_t)(im >> 1);
    nidx[2] = jp == j ? col : (jp == 0 ? col - L.ywrap : col + rs);
    nidx[3] = jm == j ? col : (jm == ny - 1 && j == 0 ? col + L.ywrap : col - rs);
    nidx[4] = kp == k ? col : (kp == 0 ? col - L.zwrap : col + pl);
    nidx[5] = km == k ? col : (km == nz - 1 && k == 0 ? col + L.zwrap : col - pl);
  }
#pragma unroll
  for (int d = 0; d < 6; d++) {
    const uint32_t q = nidx[d];
    s.nb[d] = V3{ldx(X, q), ldx(Y, q), ldx(Z, q)};
    if constexpr (MASK) {
      if constexpr (NBPT) {
        if (ldx(PT, q) == 0) ex &= ~(1u << d);
      }
      if constexpr (NBX) {
        if (d < 2) s.qx[d] = q;
      }
    }
  }
  s.ex = ex;
  s.sv = V3{0.0f, 0.0f, 0.0f};
  if constexpr (SURF) {
    if (s.ptv > 1) s.sv = V3{ldx(SX, self), ldx(SY, self), ldx(SZ, self)};
  }
  return true;
}

// gather() for a row-aligned lattice: nxh a multiple of 64 (nx a multiple of 128). The 64 positions [g, g + 64) of
// a lane group then lie in one row, so the row, j, k, the parity of the row, the y / z wraps and the y / z free
// faces are wave-uniform (scalar registers): lane `lane` is position g + lane, and only its x neighbours, the point
// types and the anchoring vectors depend on the lane. Same sites, neighbours, existence mask and order as gather(),
// missing neighbours included (they keep the site's own column of the other colour and are masked out).
template <bool MASK, bool SURF, bool FREE, bool NBX, bool NBPT = true>
__device__ __forceinline__ bool gather_row(const LatDev &L, int colour, uint32_t g, int lane,
                                           const float *__restrict__ X, const float *__restrict__ Y,
                                           const float *__restrict__ Z, const uint8_t *__restrict__ PT,
                                           const float *__restrict__ SX, const float *__restrict__ SY,
                                           const float *__restrict__ SZ, Site &s, uint32_t &self) {
  const uint32_t nxh = (uint32_t)L.nxh;
  const uint32_t row = fdiv(g, L.mxh, L.sxh);
  const uint32_t m0 = g - row * nxh;
  const uint32_t k = fdiv(row, L.my, L.sy);
  const uint32_t j = row - k * (uint32_t)L.ny;
  const uint32_t par = (j + k + (uint32_t)colour) & 1u;
  const uint32_t rowoff = L.off + row * nxh;
  const uint32_t ob = rowoff + (uint32_t)((1 - colour) * L.nc);  // other colour, this row
  const uint32_t m = m0 + (uint32_t)lane;
  self = rowoff + (uint32_t)(colour * L.nc) + m;
  s.i = (int)(2 * m + par);
  s.j = (int)j;
  s.k = (int)k;
  s.ptv = 1;
  if constexpr (MASK) {
    s.ptv = ldx(PT, self);
    if (s.ptv == 0) return false;
  }
  s.n = V3{ldx(X, self), ldx(Y, self), ldx(Z, self)};
  const uint32_t col = ob + m;
  unsigned ex = 63u;
  // parity 0 (i = 2m): +x neighbour at m, -x at m - 1; parity 1 (i = 2m + 1): +x at m + 1, -x at m. At the row's
  // ends the neighbour wraps (periodic x) or is missing (free x).
  uint32_t xp, xm;
  if (par) {
    const bool edge = m + 1 == nxh;
    xp = edge ? ob : col + 1;
    xm = col;
    if constexpr (FREE) {
      if (!(L.flags & F_PX) && edge) {
        xp = col;
        ex &= ~1u;
      }
    }
  } else {
    const bool edge = m == 0;
    xp = col;
    xm = edge ? ob + nxh - 1 : col - 1;
    if constexpr (FREE) {
      if (!(L.flags & F_PX) && edge) {
        xm = col;
        ex &= ~2u;
      }
    }
  }
  uint32_t yp = (j + 1 == (uint32_t)L.ny ? ob - L.ywrap : ob + nxh) + m;
  uint32_t ym = (j == 0 ? ob + L.ywrap : ob - nxh) + m;
  uint32_t zp = (k + 1 == (uint32_t)L.nz ? ob - L.zwrap : ob + L.plane) + m;
  uint32_t zm = (k == 0 ? ob + L.zwrap : ob - L.plane) + m;
  if constexpr (FREE) {
    if (!(L.flags & F_PY)) {
      if (j + 1 == (uint32_t)L.ny) {
        yp = col;
        ex &= ~4u;
      }
      if (j == 0) {
        ym = col;
        ex &= ~8u;
      }
    }
    if (!(L.flags & F_PZ)) {
      if (k + 1 == (uint32_t)L.nz) {
        zp = col;
        ex &= ~16u;
      }
      if (k == 0) {
        zm = col;
        ex &= ~32u;
      }
    }
  }
  const uint32_t nidx[6] = {xp, xm, yp, ym, zp, zm};
#pragma unroll
  for (int d = 0; d < 6; d++) {
    const uint32_t q = nidx[d];
    s.nb[d] = V3{ldx(X, q), ldx(Y, q), ldx(Z, q)};
    if constexpr (MASK) {
      if constexpr (NBPT) {
        if (ldx(PT, q) == 0) ex &= ~(1u << d);
      }
      if constexpr (NBX) {
        if (d < 2) s.qx[d] = q;
      }
    }
  }
  s.ex = ex;
  s.sv = V3{0.0f, 0.0f, 0.0f};
  if constexpr (SURF) {
    if (s.ptv > 1) s.sv = V3{ldx(SX, self), ldx(SY, self), ldx(SZ, self)};
  }
  return true;
}

// The site alone (no neighbours): decodes position t, loads its point type, director and anchoring vector.
template <bool MASK, bool SURF>
__device__ __forceinline__ bool gather_self(const LatDev &L, int colour, int t, const float *__restrict__ X,
                                            const float *__restrict__ Y, const float *__restrict__ Z,
                                            const uint8_t *__restrict__ PT, const float *__restrict__ SX,
                                            const float *__restrict__ SY, const float *__restrict__ SZ, Site &s) {
  const uint32_t row = fdiv((uint32_t)t, L.mxh, L.sxh);
  const int m = t - (int)row * L.nxh;
  const int k = (int)fdiv(row, L.my, L.sy);
  const int j = (int)row - k * L.ny;
  if (m < 0 || m >= L.nxh || j < 0 || j >= L.ny || k >= L.nz) return false;
  const int i = 2 * m + ((j + k + colour) & 1);
  if (i >= L.nx) return false;
  s.i = i;
  s.j = j;
  s.k = k;
  const uint32_t self = L.off + (uint32_t)(colour * L.nc) + (uint32_t)t;
  s.ptv = 1;
  if constexpr (MASK) {
    s.ptv = ldx(PT, self);
    if (s.ptv == 0) return false;
  }
  s.n = V3{ldx(X, self), ldx(Y, self), ldx(Z, self)};
  s.ex = 0u;
  s.sv = V3{0.0f, 0.0f, 0.0f};
  if constexpr (SURF) {
    if (s.ptv > 1) s.sv = V3{ldx(SX, self), ldx(SY, self), ldx(SZ, self)};
  }
  return true;
}

// One DPP move of a double (both 32-bit halves) across the lanes of a row of 16.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double x, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                          __builtin_amdgcn_readlane(__double2loint(x), l));
}

// Sum of a double over the 64 lanes of a wave, in a fixed order, without LDS: DPP butterflies inside each row of 16
// lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror 4, row_ror 8), then the four row sums read from lanes 0,
// 16, 32 and 48. Every lane gets the total. All lanes of the wave must be active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x128>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Row sums of 8 doubles at once by reduce-scatter, in a fixed order. Each step pairs lanes that hold the same
// subset of the quantities; each lane keeps half of them and adds its partner's copies: row_half_mirror (partner
// l ^ 7, lane bit 2 picks the half), quad_perm [1,0,3,2] (l ^ 1, bit 0), quad_perm [2,3,0,1] (l ^ 2, bit 1), then
// row_ror 8 (l ^ 8) adds the other half-row. Lane l ends with the sum over its row of 16 lanes of quantity
// rs8_index(l); lanes l and l ^ 8 hold the same value. About 50 VALU instructions for the 8 quantities, against 8
// wave_sum calls of about 23 each. All lanes of the wave must be active.
__device__ __forceinline__ int rs8_index(int lane) { return (lane & 4) + ((lane & 1) << 1) + ((lane & 2) >> 1); }

__device__ __forceinline__ double row_rs8(const double q[8], int lane) {
  const bool b2 = lane & 4, b0 = lane & 1, b1 = lane & 2;
  double r[4], s[2];
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = (b2 ? q[i + 4] : q[i]) + dpp_f64<0x141>(b2 ? q[i] : q[i + 4]);
#pragma unroll
  for (int j = 0; j < 2; j++) s[j] = (b0 ? r[j + 2] : r[j]) + dpp_f64<0xB1>(b0 ? r[j] : r[j + 2]);
  const double t = (b1 ? s[1] : s[0]) + dpp_f64<0x4E>(b1 ? s[0] : s[1]);
  return t + dpp_f64<0x128>(t);
}

// ang_var rule of evolve.cu:430-436. The reference compares the double quotient acc / nvalid with 0.5; for
// integers below 2^31 that is the same as comparing 2 acc with nvalid (the quotient is never within rounding of
// 0.5 unless it equals 0.5), which avoids an FP64 division. The updates keep the reference's float * double
// arithmetic.
__host__ __device__ __forceinline__ float ang_var_rule(float av, unsigned acc, int nvalid) {
  if (nvalid <= 0) return av;
  const long long twice = 2ll * (long long)acc;
  if (twice < (long long)nvalid) {
    av = (float)((double)av * 0.99);
    if ((double)av < 0.01) av = 0.01f;
  } else if (twice > (long long)nvalid) {
    av = (float)((double)av / 0.99);
    if ((double)av > 1.0) av = (float)((double)av - 0.5);
  }
  return av;
}

// ---------------------------------------------------------------------------------------------------------------
// Sweep pass over one colour of every lattice. blk[L] = (lattice, first colour-block position) of logical
// workgroup L; consecutive logical workgroups are placed on the same XCD (workgroups are dispatched to the 8 XCDs
// round robin), so the rows a workgroup shares with its neighbours in z and y mostly hit the same L2.
__device__ __forceinline__ double eig_sym3_max(const double a[6], double v[3]);

// Compact layout of the measurement partials written by the sweep itself: 0 E, 1..5 sum n_a n_b (xx xy xz yy yz),
// 6 count, 7 sum n_z^2; in contexts that measure a radial order (sphere or cylinder, one per lattice), also
// 8 sum (3 c^2 - 1) / 2, 9 sum c, 10 count of the sites off the centre (c = n.r / r, parameter_order.cpp:223-245).
constexpr int kNQS = 8;
constexpr int kNQSR = 11;
__host__ __device__ constexpr bool radial_flags(int flags) { return (flags & (F_RSPH | F_RCYL)) != 0; }

// Finalize of one lattice's compact partials [rg.x, rg.y) (column stride `stride`) by a whole workgroup of any
// multiple of 64 threads: fixed-order sum, observables, and the sample added to the moments. A lattice with a radial
// order (flags) has kNQSR quantities, its radial sums in rows 8-10; the pear observables are not in this layout.
// RAD: the lattice may carry a radial order (the generic sweep instances; a context measuring one never takes the
// row-aligned path, whose plain passes keep this finalize at its smallest).
template <bool RAD>
__device__ __forceinline__ void finalize_compact(int r, int flags, const double *__restrict__ part, int stride,
                                                 int2 rg, MomDev *__restrict__ mom, ObsDev *__restrict__ obs) {
  // The summation order of k_finalize whatever the sweep's block size: kFinWaves virtual waves of 64 lanes, lane l of
  // virtual wave vw adding partials rg.x + 64 vw + l, + kFinBlock, ...; each virtual wave summed by wave_sum; the
  // wave sums added in order. A measured sweep then gets the same moments, bit for bit, whether its finalize runs
  // here (in the next colour-0 pass) or in k_finalize (after the last sweep of a call or of a captured graph).
  // The (virtual wave, quantity) pairs are spread over the workgroup's waves, one accumulator each: the finalize is
  // inlined into every sweep instance, whose plain passes are held to 64 VGPRs. A lane loads four of its partials
  // at a time and adds them in increasing order, so the one finalizing workgroup waits for few memory round trips
  // and ends well within the colour pass it runs beside.
  __shared__ double fred[kFinWaves][RAD ? kNQSR : kNQS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  const int nq = RAD && radial_flags(flags) ? kNQSR : kNQS;
#pragma unroll 1
  for (int j = w; j < kFinWaves * nq; j += nw) {
    const int vw = !RAD || nq == kNQS ? j / kNQS : j / kNQSR, a = j - vw * nq;
    const double *__restrict__ col = part + (size_t)a * stride;
    double v = 0.0;
#pragma unroll 1
    for (int p = rg.x + 64 * vw + lane; p < rg.y; p += 4 * kFinBlock) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; u++) x[u] = p + u * kFinBlock < rg.y ? col[p + u * kFinBlock] : 0.0;
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (p + u * kFinBlock < rg.y) v += x[u];
    }
    const double x = wave_sum(v);
    if (lane == 0) fred[vw][a] = x;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t[kNQS];
#pragma unroll
  for (int a = 0; a < kNQS; a++) {
    double x = 0.0;
#pragma unroll 1
    for (int ww = 0; ww < kFinWaves; ww++) x += fred[ww][a];
    t[a] = x;
  }
  const double N = t[6], inv = N > 0.0 ? 1.0 / N : 0.0;
  ObsDev o = {};
  o.nvalid = N;
  o.E = t[0] * inv;
  o.Q[0] = (3.0 * t[1] - N) * 0.5 * inv;
  o.Q[1] = 3.0 * t[2] * 0.5 * inv;
  o.Q[2] = 3.0 * t[3] * 0.5 * inv;
  o.Q[3] = (3.0 * t[4] - N) * 0.5 * inv;
  o.Q[4] = 3.0 * t[5] * 0.5 * inv;
  o.Q[5] = (3.0 * t[7] - N) * 0.5 * inv;
  double qm[6] = {o.Q[0], o.Q[1], o.Q[2], o.Q[3], o.Q[4], -o.Q[0] - o.Q[3]};
  o.S = N > 0.0 ? eig_sym3_max(qm, o.dir) : 0.0;
  if (N <= 0.0) { o.dir[0] = 0; o.dir[1] = 0; o.dir[2] = 1; }
  if (RAD && nq > kNQS) {
    // the radial order, in k_finalize's order and arithmetic (its tot[7..9] or tot[10..12])
    double rs = 0.0, rp = 0.0, rn = 0.0;
#pragma unroll 1
    for (int ww = 0; ww < kFinWaves; ww++) {
      rs += fred[ww][8];
      rp += fred[ww][9];
      rn += fred[ww][10];
    }
    const double sr = rn > 0.0 ? rs / rn : 0.0, pr = rn > 0.0 ? rp / rn : 0.0;
    if (flags & F_RSPH) {
      o.srs = sr;
      o.prs = pr;
    } else {
      o.src = sr;
      o.prc = pr;
    }
  }
  obs[r] = o;
  MomDev m = mom[r];
  m.n += 1.0;
  m.S += o.S;
  m.S2 += o.S * o.S;
  m.E += o.E;
  m.E2 += o.E * o.E;
  if constexpr (RAD) {
    m.srs += o.srs;
    m.prs += o.prs;
    m.src += o.src;
    m.prc += o.prc;
  }
  mom[r] = m;
}

// One Metropolis trial of site s with the draws of its Philox block R (evolve.cu:162-239): rotation by pi va about x,
// y or z, renormalization, acceptance, and the optional flip trial. With two sites per thread (SPT = 2) site e of
// the pair takes half of R's 128 bits. Returns whether the rotation was accepted; nc and E_fin are the site's final
// director and energy, and changed tells whether nc differs from s.n.
template <int POT, bool SURF, bool FIELD, bool FLIP, int SPT>
__device__ __forceinline__ bool metropolis_trial(const LatDev &L, const Site &s, const mcl::Philox4 &R, int e, float av,
                                                 float T, float invT, float elecE,
                                                 const SurfDev *__restrict__ surf, V3 &nc, float &E_fin,
                                                 bool &changed) {
  float u_ang, rot, u;
  if constexpr (SPT == 2) {
    // selects, not R.v[2 * e]: a run-time e would index the private array, which the compiler then moves to LDS
    // (16 B per thread of the 1024-thread bound) and reads back there
    const uint32_t w = e ? R.v[2] : R.v[0];
    u_ang = mcl::u01_open16(w >> 16);
    rot = mcl::u01_open16(w & 0xFFFFu);
    u = mcl::u01_open(e ? R.v[3] : R.v[1]);
  } else {
    u_ang = mcl::u01_open(R.v[0]);
    rot = mcl::u01_open(R.v[1]);
    u = mcl::u01_open(R.v[2]);
  }
  // trial rotation by pi * va about x, y or z (evolve.cu:162-182)
  const float va = (2.0f * u_ang - 1.0f) * av;
  // cos and sin of pi va: the hardware sine and cosine take their argument in turns (x 2 pi)
  const float sn = __builtin_amdgcn_sinf(0.5f * va), cs = __builtin_amdgcn_cosf(0.5f * va);
  const V3 n = s.n;
  V3 nn;
  if (rot < 0.333f) {
    nn = V3{n.x, fmaf(n.y, cs, n.z * sn), fmaf(n.z, cs, -(n.y * sn))};
  } else if (rot < 0.666f) {
    nn = V3{fmaf(n.x, cs, -(n.z * sn)), n.y, fmaf(n.z, cs, n.x * sn)};
  } else {
    nn = V3{fmaf(n.x, cs, n.y * sn), fmaf(n.y, cs, -(n.x * sn)), n.z};
  }
  const float inv = __builtin_amdgcn_rsqf(dot3(nn, nn));  // |nn| ~ 1: no denormal handling needed
  nn.x *= inv;
  nn.y *= inv;
  nn.z *= inv;
  const float E_old = site_energy<POT, SURF, FIELD>(L, s, n, surf, elecE);
  const float E_new = site_energy<POT, SURF, FIELD>(L, s, nn, surf, elecE);
  // Metropolis; T = 0 accepts when the energy does not increase (evolve.cu:212-224)
  const bool accepted = T > 0.0f ? (u <= __expf(-(E_new - E_old) * invT)) : (E_new <= E_old);
  nc = accepted ? nn : n;
  E_fin = accepted ? E_new : E_old;
  changed = accepted;
  if constexpr (FLIP) {
    // separate n -> -n trial, not counted in the acceptance (evolve.cu:225-239)
    const V3 nf{-nc.x, -nc.y, -nc.z};
    const float E_f = site_energy<POT, SURF, FIELD>(L, s, nf, surf, elecE);
    const float uf = mcl::u01_open(R.v[3]);
    if (T > 0.0f ? (uf <= __expf(-(E_f - E_fin) * invT)) : (E_f <= E_fin)) {
      nc = nf;
      E_fin = E_f;
      changed = true;
    }
  }
  return accepted;
}

// EMPTY = false (masked GHRL and pear contexts none of whose lattices has an empty site, such as the slab): the
// neighbours' point types are not loaded, since none of them can be 0 (LL never loads them, see gather's NBPT).
// SPLIT (small contexts without the flip trial, not row-aligned): each thread takes one position instead of a pair,
// computes the pair's Philox draw itself and uses the half the pair's thread would give that position, so the
// trials are the same and a wave's serial work halves (few waves per SIMD; LL lattices up to 80^3).
template <int POT, bool MASK, bool SURF, bool FIELD, bool FLIP, bool FREE, bool MEAS, bool ROW, bool EMPTY = true,
          bool SPLIT = false, bool XR = false>
// Launch shape (chosen by the autotuner): any multiple of 64 up to 1024 threads per workgroup; with xcd != 0
// workgroup b runs logical tile (b & 7) * chunk + (b >> 3), so consecutive tiles share an XCD and its L2.
// The sweep number is *sweep_base + k: graphs replay fixed arguments while the base advances on the device.
// MEAS (colour-1 pass of a measured sweep, unmasked lattices with at least two sites along x): the pass also
// measures the lattice it leaves behind. Every bond has exactly one colour-1 end and every colour-1 site's final
// energy is known from its trial, so the lattice energy is the sum over colour-1 sites of twice their final site
// energy minus their one-site terms, plus the one-site terms of colour 0. Each colour-0 site is counted (Q tensor,
// one-site terms, count) by the colour-1 site on its +x side, or, at the last column of a free x axis, by the one on
// its -x side. The workgroup writes kNQS FP64 partials to spart[q * nblk + workgroup] (kNQSR with XR, which adds the
// radial order of the lattices that measure one). In an LL context with empty sites (stored as zero vectors) an
// empty colour-1 position still counts the colour-0 sites it owns, and an empty colour-0 site is not counted.
// ROW: every lattice of the context is row-aligned (gather_row): site decoding, the y / z neighbour indices and
// the y / z free faces are scalar work.
// A launch sweeps the workgroups [b0, b1) of the work list (one lattice group, or all of it); nblk is the length of
// the whole list (the column stride of spart).
// Register bound: 8 waves per SIMD (64 VGPRs); 6 for the measuring passes of masked, radial or non-LL contexts, 5 for
// those that are masked and radial (at 6 they spilled 20-24 bytes per lane). No instance uses scratch.
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(
    MEAS && MASK && XR ? 5 : (MEAS && (MASK || XR || POT != MCL_POT_LL) ? 6 : 8), 8))) void k_sweep(
    const LatDev *__restrict__ lats, CtrlDev *__restrict__ ctrl,
                                                unsigned int *__restrict__ cnt, const int2 *__restrict__ blk,
                                                int nblk, int b0, int b1, int chunk, int xcd, int colour,
                                                const unsigned long long *__restrict__ sweep_base, unsigned k,
                                                  float *__restrict__ X, float *__restrict__ Y,
                                                  float *__restrict__ Z, const uint8_t *__restrict__ PT,
                                                  const float *__restrict__ SX, const float *__restrict__ SY,
                                                  const float *__restrict__ SZ, const SurfDev *__restrict__ surf,
                                                  double *__restrict__ spart, int fin, int l0, int nfin,
                                                  const int2 *__restrict__ srange, MomDev *__restrict__ mom,
                                                  ObsDev *__restrict__ obs) {
  const int b = blockIdx.x;
  // logical workgroups [0, nfin) finalize the measurement folded into the previous sweep (fin = 1), one per lattice
  // l0 + L0: they come first in dispatch order, so their FP64 eigen-solve runs beside the sweep tiles instead of
  // after them. nfin is the number of lattices of the launch when the context measures in the sweep, else 0.
  const int L0 = xcd ? (b & 7) * chunk + (b >> 3) : b;
  if (L0 < nfin) {
    if (fin) finalize_compact<!ROW>(l0 + L0, ROW ? 0 : lats[l0 + L0].flags, spart, nblk, srange[l0 + L0], mom, obs);
    return;
  }
  const int Lb = b0 + L0 - nfin;
  if (Lb >= b1) return;
  const int2 bd = blk[Lb];
  const int r = bd.x;
  const LatDev &L = lats[r];
  const unsigned long long sweep = *sweep_base + k;
  const float T = ctrl[r].T, invT = ctrl[r].invT, elecE = ctrl[r].elecE;
  const int lane = threadIdx.x & 63;
  static_assert(!SPLIT || (!FLIP && !ROW), "SPLIT is for pair-sharing generic passes");
  const int lblk = bd.y / ((FLIP || SPLIT ? 1 : 2) * (int)blockDim.x);  // workgroup index within the lattice
  unsigned int *const cnt_r = cnt + (size_t)r * kCntLattice;
  // Sites per thread: without the flip trial a pair of positions (p, p + 64) of the colour block shares one
  // Philox4x32-10 draw, 64 bits per site (16-bit angle, 16-bit axis choice, 32-bit acceptance uniform): each wave
  // updates 128 consecutive positions, and each of its loads stays one contiguous 256-byte span. With the flip
  // trial every site takes a draw of its own (a fourth 32-bit uniform). The counter (pair or position, sweep,
  // tag + colour) depends on the position only, never on the launch shape (a workgroup spans a multiple of 128
  // positions).
  constexpr int SPT = FLIP ? 1 : 2;
  const int tw = SPT == 2 && !SPLIT ? bd.y + 128 * (int)(threadIdx.x >> 6) + lane : bd.y + (int)threadIdx.x;
  const uint32_t ctr0 = SPLIT ? (uint32_t)((tw >> 7) * 64 + (tw & 63))
                              : (SPT == 2 ? (uint32_t)(((bd.y >> 7) + (int)(threadIdx.x >> 6)) * 64 + lane)
                                          : (uint32_t)tw);
  const mcl::Philox4 R = mcl::philox4x32_10(ctr0, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                            kTagSweep + (uint32_t)colour, L.key0, L.key1);
  // ang_var. Colour 1 reads what the lattice's first colour-0 workgroup published. In a colour-0 pass the accepted
  // rotations of the previous sweep are summed from the lattice's counters (kCnt of them, L.kc in masked contexts
  // that are not row-aligned; one counter, one cache line, per lane) and the rule applied; the counter load is issued
  // before the check for an ang_var set by the host, so that it does not wait for the control block.
  // W0 (unmasked contexts): only the first wave of each workgroup reads the counters; the other waves issue their
  // first gather and take ang_var from LDS after a workgroup barrier placed behind it (take_av in the trial loop).
  // 128^3 +3.3 % (8 times fewer requests to the 64 counter lines); masked contexts keep one read per wave (the slab
  // was 2.5 % slower with the barrier). Measured alternatives, all slower at 128^3: wave-uniform (scalar) loads of
  // the counters, 16 %; the first wave reading them and passing ang_var through LDS after a barrier before the
  // gather, 1.3 %; every wave's sum and rule after its first gather, 4.5 %, and with the counter load itself issued
  // after the gather, 3 %; a DPP sum instead of the LDS permutes, 0.6 %; 16 or 8 counters, 1.4 or 2.8 % (their
  // atomics queue). Passing the work list and the sweep base as kernel arguments to eager launches of one-lattice
  // contexts (a runtime switch, or an instance of its own) gave -0.6 to -1.5 % at 128^3.
  constexpr bool W0 = !MASK;
  float av = 0.0f;
  __shared__ float s_av;
  if (colour == 1) {
    av = ctrl[r].av_store[sweep & 1];
  } else if (!W0 || __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
    unsigned a = lane < (MASK && !ROW ? L.kc : kCnt) ? cnt_r[((sweep - 1) % 3) * kCntSweep + lane * kCntStride] : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    const float ruled =
        ang_var_rule(ctrl[r].av_store[(sweep - 1) & 1], (unsigned)__builtin_amdgcn_readfirstlane(a), L.nvalid);
    const float avw = sweep == ctrl[r].av_set_sweep ? ctrl[r].av_set : ruled;
    if constexpr (W0) {
      if (lane == 0) s_av = avw;
    } else {
      av = avw;
    }
  }
  // the lattice's first workgroup publishes ang_var for colour 1 and the next sweep, and clears the counters of
  // sweep + 1 (last read by the colour-0 pass of sweep - 1)
  auto publish = [&]() {
    if (lblk == 0 && threadIdx.x < kCnt) {
      if (threadIdx.x == 0) ctrl[r].av_store[sweep & 1] = av;
      cnt_r[((sweep + 1) % 3) * kCntSweep + threadIdx.x * kCntStride] = 0u;
    }
  };
  if (!W0 && colour == 0) publish();
  // W0: after the first gather of every wave (all waves of the workgroup reach it)
  auto take_av = [&]() {
    if constexpr (W0) {
      if (colour == 0) {
        __syncthreads();
        av = s_av;
        publish();
      }
    }
  };
  unsigned nacc = 0;
  static_assert(!XR || MEAS, "XR is a measuring instance");
  constexpr int NQS = XR ? kNQSR : kNQS;
  double qm[NQS];
#pragma unroll
  for (int a = 0; a < NQS; a++) qm[a] = 0.0;
  auto add_q = [&](V3 v) {
    const double x = v.x, y = v.y, z = v.z;
    qm[1] += x * x;
    qm[2] += x * y;
    qm[3] += x * z;
    qm[4] += y * y;
    qm[5] += y * z;
    qm[7] += z * z;
    qm[6] += 1.0;
  };
  // radial order of a site at (i, j, k): c = n.r / r about the lattice's centre, sites within half a spacing of it
  // (or of the axis) skipped; r in the xy plane for a cylinder (parameter_order.cpp:223-245, as k_measure)
  auto add_r = [&](V3 v, int i, int j, int k) {
    if constexpr (XR) {
      if (radial_flags(L.flags)) {
        const float rx = i - L.cx, ry = j - L.cy, rz = (L.flags & F_RSPH) ? k - L.cz : 0.0f;
        const float rr = sqrtf(rx * rx + ry * ry + rz * rz);
        if (rr >= 0.5f) {
          const double c = (double)((v.x * rx + v.y * ry + v.z * rz) / rr);
          qm[8] += 1.5 * c * c - 0.5;
          qm[9] += c;
          qm[10] += 1.0;
        }
      }
    }
  };
  // an empty colour-1 position t of an LL context (its director is a zero vector, no trial) still counts the
  // colour-0 sites it owns: its -x neighbour, and its +x neighbour at the end of a free x axis
  auto own_from_empty = [&](uint32_t t) {
    if constexpr (MEAS && MASK && POT == MCL_POT_LL) {
      const uint32_t row = fdiv(t, L.mxh, L.sxh);
      const int m = (int)(t - row * (uint32_t)L.nxh);
      const int k = (int)fdiv(row, L.my, L.sy);
      const int j = (int)row - k * L.ny;
      const int i = 2 * m + ((j + k + 1) & 1);
      if (i >= L.nx) return;  // padding position of an odd x axis
      const uint32_t rowstart = L.off + row * (uint32_t)L.nxh;  // this row of the colour-0 block
      double e1 = 0.0;
      auto own0 = [&](int ii) {
        const uint32_t q = rowstart + (uint32_t)(ii >> 1);
        const int pt = ldx(PT, q);
        if (pt == 0) return;
        const V3 v{ldx(X, q), ldx(Y, q), ldx(Z, q)};
        V3 sv{0.0f, 0.0f, 0.0f};
        if constexpr (SURF) {
          if (pt > 1) sv = V3{ldx(SX, q), ldx(SY, q), ldx(SZ, q)};
        }
        add_q(v);
        if constexpr (SURF || FIELD) e1 += (double)site_terms<SURF, FIELD>(L, pt, sv, v, surf, elecE);
        add_r(v, ii, j, k);
      };
      if (i > 0)
        own0(i - 1);
      else if (L.flags & F_PX)
        own0(L.nx - 1);
      if (!(L.flags & F_PX) && i + 1 == L.nx - 1) own0(i + 1);
      qm[0] += e1;
    }
  };
#pragma unroll
  for (int e = 0; e < (SPLIT ? 1 : SPT); e++) {
    Site s;
    uint32_t self = 0;
    if constexpr (ROW) {
      // first position of this lane group, wave-uniform (a wave covers 64 SPT consecutive positions)
      const uint32_t g = (uint32_t)bd.y + (uint32_t)(64 * SPT * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) +
                         (uint32_t)(64 * e);
      const bool in = g < (uint32_t)L.nc;
      const bool ok = in && gather_row<MASK, SURF, FREE, MEAS && MASK, EMPTY && POT != MCL_POT_LL>(
                                L, colour, g, lane, X, Y, Z, PT, SX, SY, SZ, s, self);
      if (e == 0) take_av();
      if (!ok) {
        if (in) own_from_empty(g + (uint32_t)lane);
        continue;
      }
    } else {
      const int t = SPLIT ? tw : tw + 64 * e;
      const bool in = t < L.nc;
      const bool ok = in && gather<MASK, SURF, FREE, MEAS && MASK, EMPTY && POT != MCL_POT_LL>(
                                L, colour, t, X, Y, Z, PT, SX, SY, SZ, s, self);
      if (e == 0) take_av();
      if (!ok) {
        if (in) own_from_empty((uint32_t)t);
        continue;
      }
    }
    V3 nc;
    float E_fin;
    bool changed;
    nacc += metropolis_trial<POT, SURF, FIELD, FLIP, SPT>(L, s, R, SPLIT ? (tw >> 6) & 1 : e, av, T, invT, elecE,
                                                          surf, nc, E_fin, changed)
                ? 1u
                : 0u;
    // Every swept site is stored, changed or not (nc holds the old bits after a rejected trial), so each wave
    // writes whole 256-byte spans per component. Storing only the changed sites leaves about half of the bytes of
    // every line masked; at 128^3 that pass took 8.59 us against 8.22 us.
    (void)changed;
    stx(X, self, nc.x);
    stx(Y, self, nc.y);
    stx(Z, self, nc.z);
    if constexpr (MEAS) {
      // one-site terms (anchoring on surface sites, field) of a site of point type pt and anchoring vector sv
      auto terms = [&](int pt, V3 sv, V3 v) -> double {
        if constexpr (SURF || FIELD) return (double)site_terms<SURF, FIELD>(L, pt, sv, v, surf, elecE);
        return 0.0;
      };
      double e2 = 2.0 * (double)E_fin;
      add_q(nc);
      add_r(nc, s.i, s.j, s.k);
      e2 -= terms(s.ptv, s.sv, nc);
      // a colour-0 site this site counts (x neighbour d, at x position ii): Q tensor, one-site terms, radial order.
      // In a masked lattice its point type and anchoring vector are loaded here; an empty one (LL contexts) is not
      // counted.
      auto own = [&](int d, int ii) {
        int pt = 1;
        V3 sv{0.0f, 0.0f, 0.0f};
        if constexpr (MASK) {
          const uint32_t q = s.qx[d];
          pt = ldx(PT, q);
          if constexpr (POT == MCL_POT_LL) {
            if (pt == 0) return;
          }
          if constexpr (SURF) {
            if (pt > 1) sv = V3{ldx(SX, q), ldx(SY, q), ldx(SZ, q)};
          }
        }
        add_q(s.nb[d]);
        e2 += terms(pt, sv, s.nb[d]);
        add_r(s.nb[d], ii, s.j, s.k);
      };
      // the colour-0 sites this site counts: its -x neighbour, and its +x neighbour at the end of a free x axis
      const bool own_m = FREE ? (s.ex & 2u) != 0 : true;
      const bool own_p = FREE && !(L.flags & F_PX) && s.i + 1 == L.nx - 1;
      if (own_m) own(1, s.i > 0 ? s.i - 1 : L.nx - 1);
      if (own_p) own(0, s.i + 1);
      qm[0] += e2;
    }
  }

  // accepted rotations: wave ballots, then one atomic per workgroup into one of kCnt counters (striping keeps the
  // device-scope atomics of thousands of workgroups from serializing on one address)
  __shared__ unsigned int wsum[16];
  unsigned wc = (unsigned)__popcll(__ballot(nacc & 1u));
  if constexpr (SPT == 2) wc += 2u * (unsigned)__popcll(__ballot(nacc >> 1));
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = wc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 0; w < nw; w++) tot += wsum[w];
    if (tot) atomicAdd(&cnt_r[(sweep % 3) * kCntSweep + (lblk & ((MASK && !ROW ? L.kc : kCnt) - 1)) * kCntStride], tot);
  }
  if constexpr (MEAS && (MASK || XR || POT != MCL_POT_LL)) {
    static_assert(kNQS == 8, "row_rs8 reduces 8 quantities");
    // per wave and row of 16 lanes, the row sums of the 8 quantities (lanes 0-7 of each row write them). Only in
    // the instances that run at 6 or 5 waves per SIMD: under the 8-wave bound of the others the reduce-scatter
    // spills, and they reduce through LDS (below).
    // The measuring passes that get here are those of GHRL contexts, of masked LL contexts (slabs, spheres, lattices
    // with empty sites) and of contexts that measure a radial order; pear contexts measure with the separate
    // kernels.
    __shared__ double mred[16][4][kNQS];
    const double x = row_rs8(qm, lane);
    if ((lane & 8) == 0) mred[threadIdx.x >> 6][lane >> 4][rs8_index(lane)] = x;
    // the radial sums (XR) by whole-wave sums
    __shared__ double xred[XR ? 16 : 1][XR ? kNQSR - kNQS : 1];
    if constexpr (XR) {
#pragma unroll
      for (int a = kNQS; a < NQS; a++) {
        const double v = wave_sum(qm[a]);
        if (lane == 0) xred[threadIdx.x >> 6][a - kNQS] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x < kNQS) {
      double v = 0.0;
      const int nw = (int)(blockDim.x >> 6);
      for (int w = 0; w < nw; w++)
        v += (mred[w][0][threadIdx.x] + mred[w][1][threadIdx.x]) + (mred[w][2][threadIdx.x] + mred[w][3][threadIdx.x]);
      spart[(size_t)threadIdx.x * nblk + Lb] = v;
    } else if (XR && threadIdx.x < NQS) {
      double v = 0.0;
      const int nw = (int)(blockDim.x >> 6);
      for (int w = 0; w < nw; w++) v += xred[w][threadIdx.x - kNQS];
      spart[(size_t)threadIdx.x * nblk + Lb] = v;
    }
  } else if constexpr (MEAS) {
    // unmasked LL contexts (8 waves per SIMD): every thread leaves its sums in the dynamic LDS, then wave w adds
    // quantity w over the workgroup's threads in order (lane by lane, wave after wave) and sums its lanes: about 35
    // VALU instructions per wave instead of kNQS wave sums (about 184 of the 615 at 128^3). Up to 512 threads all
    // kNQS quantities go at once (32 KiB at 512 threads), above that half of them at a time (32 KiB at 1024).
    extern __shared__ double mlds[];
    const int nt = (int)blockDim.x, nw = nt >> 6, w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    auto reduce = [&](int a0, int na) {
#pragma unroll 1
      for (int a = w; a < na; a += nw) {
        double v = 0.0;
#pragma unroll 1
        for (int j = 0; j < nw; j++) v += mlds[a * nt + 64 * j + lane];
        v = wave_sum(v);
        if (lane == 0) spart[(size_t)(a0 + a) * nblk + Lb] = v;
      }
    };
    if (nt <= 512) {
#pragma unroll
      for (int a = 0; a < kNQS; a++) mlds[a * nt + (int)threadIdx.x] = qm[a];
      __syncthreads();
      reduce(0, kNQS);
    } else {
      constexpr int H = kNQS / 2;
#pragma unroll
      for (int a = 0; a < H; a++) mlds[a * nt + (int)threadIdx.x] = qm[a];
      __syncthreads();
      reduce(0, H);
      __syncthreads();
#pragma unroll
      for (int a = 0; a < H; a++) mlds[a * nt + (int)threadIdx.x] = qm[H + a];
      __syncthreads();
      reduce(H, H);
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Resident sweeps of small lattices: one workgroup per lattice holds the lattice's two colour blocks in LDS and runs
// nsw whole sweeps in one launch, with a workgroup barrier between the colour passes instead of a kernel boundary.
// A small lattice is otherwise bound by the latency of two launches per sweep (a 10 x 10 x 16 slab: 16 us per
// sweep). Every trial is the one k_sweep makes: the same gather (on lattice-local positions), the same Philox
// counters (which depend on the position only), metropolis_trial, and the same ang_var rule on the accepted count
// of the previous sweep, so the directors follow k_sweep's trajectory bit for bit. At the end the lattice's
// control block and counter ring hold what k_sweep's passes would have left: ang_var of the last two sweeps, their
// accepted counts, and cleared counters for the next sweep. Plain sweeps only (measured sweeps take k_sweep).
// One CU then does all of a lattice's work, so this pays only while no lane updates more sites than in k_sweep: at
// most 2048 positions per colour (two sites per lane), 1024 with the flip trial (one). Measured on the MI355X: 1.3x
// for a 10 x 10 x 16 GHRL slab and 1.4x for LL 16^3, but 0.6-0.7x for 12 x 12 x 46 and 22^3 (4-6 sites per lane).
constexpr int kResBlock = 1024;
constexpr int kResMax = 4096;     // colour-block positions (both colours) of a lattice: 3 x 4 bytes each in LDS
constexpr int kResMaxFlip = 2048;  // the same with the flip trial (one site per lane)
constexpr int kResChunk = 16384;  // sweeps per resident launch

// ONE (no flip trial, every lattice at most kResBlock positions per colour): each thread takes one position and
// computes the draw of its pair (p, p + 64) itself, the half k_sweep gives that position, so no lane runs two
// trials in a row and twice the waves hide the latency (a separate instance: both ways in one kernel spill).
// EMPTY = false: as k_sweep's (no lattice of the context has an empty site; the neighbours' point types are not read).
template <int POT, bool MASK, bool SURF, bool FIELD, bool FLIP, bool FREE, bool ONE, bool EMPTY = true>
__global__ __launch_bounds__(kResBlock) void k_sweep_resident(
    const LatDev *__restrict__ lats, CtrlDev *__restrict__ ctrl, unsigned int *__restrict__ cnt,
    const unsigned long long *__restrict__ sweep_base, unsigned k0, int nsw, float *__restrict__ X,
    float *__restrict__ Y, float *__restrict__ Z, const uint8_t *__restrict__ PT, const float *__restrict__ SX,
    const float *__restrict__ SY, const float *__restrict__ SZ, const SurfDev *__restrict__ surf) {
  __shared__ float sd[3 * kResMax];
  __shared__ unsigned acc[2];
  const int r = blockIdx.x;
  LatDev L = lats[r];
  const uint32_t off = L.off;
  L.off = 0;  // gather on lattice-local positions: colour 0 at [0, nc), colour 1 at [nc, 2 nc)
  const int n2 = 2 * L.nc;
  float *const LX = sd, *const LY = sd + kResMax, *const LZ = sd + 2 * kResMax;
  for (int i = (int)threadIdx.x; i < n2; i += kResBlock) {
    LX[i] = X[off + i];
    LY[i] = Y[off + i];
    LZ[i] = Z[off + i];
  }
  const uint8_t *const PTl = MASK ? PT + off : nullptr;
  const float *const SXl = SURF ? SX + off : nullptr;
  const float *const SYl = SURF ? SY + off : nullptr;
  const float *const SZl = SURF ? SZ + off : nullptr;
  const float T = ctrl[r].T, invT = ctrl[r].invT, elecE = ctrl[r].elecE;
  const float av_set = ctrl[r].av_set;
  const unsigned long long av_set_sweep = ctrl[r].av_set_sweep;
  const unsigned long long sw0 = *sweep_base + k0;
  unsigned int *const cnt_r = cnt + (size_t)r * kCntLattice;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the previous sweep: its ang_var and, when the rule needs it, its accepted count (summed over the 64 counters)
  float av_prev = ctrl[r].av_store[(sw0 - 1) & 1];
  unsigned c_prev = 0;
  if (sw0 != av_set_sweep) {
    unsigned a = lane < kCnt ? cnt_r[((sw0 - 1) % 3) * kCntSweep + lane * kCntStride] : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    c_prev = (unsigned)__builtin_amdgcn_readfirstlane(a);
  }
  if (threadIdx.x < 2) acc[threadIdx.x] = 0u;
  __syncthreads();
  float av_m1 = av_prev;  // ang_var and accepted count of the sweep before the last one run here
  unsigned c_m1 = c_prev;
  for (int k = 0; k < nsw; k++) {
    const unsigned long long sweep = sw0 + (unsigned long long)k;
    const float av = sweep == av_set_sweep ? av_set : ang_var_rule(av_prev, c_prev, L.nvalid);
    for (int colour = 0; colour < 2; colour++) {
      unsigned nacc = 0;
      constexpr int SPT = FLIP ? 1 : 2;
      if constexpr (ONE) {
        const int t = (int)threadIdx.x;
        Site s;
        uint32_t self = 0;
        if (t < L.nc &&
            gather<MASK, SURF, FREE, false, EMPTY && POT != MCL_POT_LL>(L, colour, t, LX, LY, LZ, PTl, SXl, SYl, SZl, s, self)) {
          const uint32_t ctr0 = (uint32_t)((t >> 7) * 64 + (t & 63));
          const mcl::Philox4 R = mcl::philox4x32_10(ctr0, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                                    kTagSweep + (uint32_t)colour, L.key0, L.key1);
          V3 nc;
          float E_fin;
          bool changed;
          nacc = metropolis_trial<POT, SURF, FIELD, FLIP, SPT>(L, s, R, (t >> 6) & 1, av, T, invT, elecE, surf, nc,
                                                               E_fin, changed)
                     ? 1u
                     : 0u;
          if (changed) {
            LX[self] = nc.x;
            LY[self] = nc.y;
            LZ[self] = nc.z;
          }
        }
      } else {
        // positions [base, base + 64 SPT) per wave and step, as the waves of k_sweep take them
        for (int base = 64 * SPT * wave; base < L.nc; base += 64 * SPT * (kResBlock / 64)) {
          const uint32_t ctr0 = SPT == 2 ? (uint32_t)((base >> 7) * 64 + lane) : (uint32_t)(base + lane);
          const mcl::Philox4 R = mcl::philox4x32_10(ctr0, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                                    kTagSweep + (uint32_t)colour, L.key0, L.key1);
#pragma unroll
          for (int e = 0; e < SPT; e++) {
            const int t = base + lane + 64 * e;
            Site s;
            uint32_t self = 0;
            if (t >= L.nc ||
                !gather<MASK, SURF, FREE, false, EMPTY && POT != MCL_POT_LL>(L, colour, t, LX, LY, LZ, PTl, SXl, SYl, SZl, s,
                                                                     self))
              continue;
            V3 nc;
            float E_fin;
            bool changed;
            nacc += metropolis_trial<POT, SURF, FIELD, FLIP, SPT>(L, s, R, e, av, T, invT, elecE, surf, nc, E_fin,
                                                                  changed)
                        ? 1u
                        : 0u;
            if (changed) {
              LX[self] = nc.x;
              LY[self] = nc.y;
              LZ[self] = nc.z;
            }
          }
        }
      }
      unsigned wc = 0;
      for (int bit = 0; bit < 16; bit++) wc += (unsigned)__popcll(__ballot((nacc >> bit) & 1u)) << bit;
      if (lane == 0 && wc) atomicAdd(&acc[k & 1], wc);
      __syncthreads();
      // acc[(k + 1) & 1] was last read at the end of sweep k - 1, before the barrier above; the next barrier
      // orders this reset before the atomics of sweep k + 1
      if (colour == 0 && threadIdx.x == 0) acc[(k + 1) & 1] = 0u;
    }
    av_m1 = av_prev;
    c_m1 = c_prev;
    av_prev = av;
    c_prev = acc[k & 1];
  }
  for (int i = (int)threadIdx.x; i < n2; i += kResBlock) {
    X[off + i] = LX[i];
    Y[off + i] = LY[i];
    Z[off + i] = LZ[i];
  }
  if (nsw > 0 && threadIdx.x < kCnt) {
    const unsigned long long last = sw0 + (unsigned long long)(nsw - 1);
    const uint32_t q = threadIdx.x * kCntStride;
    if (threadIdx.x == 0) {
      ctrl[r].av_store[last & 1] = av_prev;
      if (nsw >= 2) ctrl[r].av_store[(last - 1) & 1] = av_m1;
    }
    cnt_r[(last % 3) * kCntSweep + q] = threadIdx.x == 0 ? c_prev : 0u;
    if (nsw >= 2) cnt_r[((last - 1) % 3) * kCntSweep + q] = threadIdx.x == 0 ? c_m1 : 0u;
    cnt_r[((last + 1) % 3) * kCntSweep + q] = 0u;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Measurement: per-workgroup FP64 partial sums over up to kBlock * kMeasPerThread sites of one colour of one lattice.
// Quantities: 0 E, 1..5 sum n_a n_b (xx xy xz yy yz), 6 nvalid, 7..9 radial sphere (sum Sr, sum c, count),
// 10..12 radial cylinder, 13..15 sum n, 16..18 splay vector, 19 bonds, 20 sum div n, 21 sum n_z^2.

// The basic quantities (energy, Q sums, count) are 0..6 and 21; XQ adds the radial, polar and pear ones.
__host__ __device__ constexpr bool basic_q(int a) { return a <= 6 || a == 21; }
// The quantities a measurement of potential POT writes: the pear observables (13..20) only for pear contexts.
template <int POT, bool XQ>
__host__ __device__ constexpr bool written_q(int a) {
  return (XQ || basic_q(a)) && (POT == MCL_POT_PEAR || a < 13 || a > 20);
}

// The energy of a lattice counts every bond twice (once from each end, as the reference's per-site sum does) plus
// the one-site terms of every site. Every bond has exactly one colour-0 end, so the colour-0 workgroups add twice
// their sites' bond sums and the colour-1 workgroups only the one-site terms, which need no neighbours (the pear
// observables, which do, gather them on both colours).
template <int POT, bool MASK, bool SURF, bool FIELD, bool XQ, bool FREE>
__global__ __launch_bounds__(kBlock) void k_measure(const LatDev *__restrict__ lats, const CtrlDev *__restrict__ ctrl,
                                                    const int4 *__restrict__ mblk, int nmblk, int m0,
                                                    const float *__restrict__ X, const float *__restrict__ Y,
                                                    const float *__restrict__ Z, const uint8_t *__restrict__ PT,
                                                    const float *__restrict__ SX, const float *__restrict__ SY,
                                                    const float *__restrict__ SZ, const SurfDev *__restrict__ surf,
                                                    double *__restrict__ partials) {
  // partials are stored quantity-major, partials[a * nmblk + workgroup], so the finalize reads them coalesced; a
  // launch measures the workgroups m0, m0 + 1, ... of the list (one lattice group, or all of it)
  const int Lb = m0 + (int)blockIdx.x;
  if (Lb >= nmblk) return;
  const int4 md = mblk[Lb];  // (lattice, colour, first position, last position + 1)
  const int r = md.x, colour = md.y;
  const LatDev &L = lats[r];
  const float elecE = ctrl[r].elecE;
  double q[kNQ];
#pragma unroll
  for (int a = 0; a < kNQ; a++) q[a] = 0.0;
  const bool bonds = colour == 0;
  constexpr bool kAllNeighbours = XQ && POT == MCL_POT_PEAR;
  for (int it = 0; it < kMeasPerThread; it++) {
    const int t = md.z + it * kBlock + (int)threadIdx.x;
    Site s;
    uint32_t self;
    if (t >= md.w) continue;
    if (bonds || kAllNeighbours) {
      if (!gather<MASK, SURF, FREE>(L, colour, t, X, Y, Z, PT, SX, SY, SZ, s, self)) continue;
    } else {
      if (!gather_self<MASK, SURF>(L, colour, t, X, Y, Z, PT, SX, SY, SZ, s)) continue;
    }
    const V3 n = s.n;
    double e = bonds ? 2.0 * (double)bond_sum<POT>(L, s, n) : 0.0;
    if constexpr (SURF || FIELD) e += (double)site_terms<SURF, FIELD>(L, s.ptv, s.sv, n, surf, elecE);
    q[0] += e;
    const double x = n.x, y = n.y, z = n.z;
    q[1] += x * x;
    q[2] += x * y;
    q[3] += x * z;
    q[4] += y * y;
    q[5] += y * z;
    q[21] += z * z;
    q[6] += 1.0;
    if (XQ && (L.flags & (F_RSPH | F_RCYL))) {
      const float rx = s.i - L.cx, ry = s.j - L.cy, rz = s.k - L.cz;
      if (L.flags & F_RSPH) {
        const float rr = sqrtf(rx * rx + ry * ry + rz * rz);
        if (rr >= 0.5f) {
          // c = n.r / r in FP32 (a float quotient widened to double), as parameter_order.cpp:236 computes it
          const double c = (double)((n.x * rx + n.y * ry + n.z * rz) / rr);
          q[7] += 1.5 * c * c - 0.5;
          q[8] += c;
          q[9] += 1.0;
        }
      }
      if (L.flags & F_RCYL) {
        const float rr = sqrtf(rx * rx + ry * ry);
        if (rr >= 0.5f) {
          const double c = (double)((n.x * rx + n.y * ry) / rr);
          q[10] += 1.5 * c * c - 0.5;
          q[11] += c;
          q[12] += 1.0;
        }
      }
    }
    if constexpr (XQ && POT == MCL_POT_PEAR) {
      q[13] += x;
      q[14] += y;
      q[15] += z;
      // splay vector over bonds (Splay_magnitude, parameter_order.cpp:147-176)
#pragma unroll
      for (int d = 0; d < 6; d++) {
        if (!(s.ex & (1u << d))) continue;
        const V3 b = s.nb[d];
        const float sg = (d & 1) ? -1.0f : 1.0f;
        const int ax = d >> 1;
        const float ra = sg * (ax == 0 ? n.x : (ax == 1 ? n.y : n.z));
        const float rb = sg * (ax == 0 ? b.x : (ax == 1 ? b.y : b.z));
        const float ab = dot3(n, b);
        q[16] += 0.5 * (b.x * rb - n.x * ra + n.x * ab * rb - b.x * ab * ra);
        q[17] += 0.5 * (b.y * rb - n.y * ra + n.y * ab * rb - b.y * ab * ra);
        q[18] += 0.5 * (b.z * rb - n.z * ra + n.z * ab * rb - b.z * ab * ra);
        q[19] += 1.0;
      }
      // divergence: central / one-sided differences per axis (Splay_divergence, parameter_order.cpp:183-215)
      double div = 0.0;
#pragma unroll
      for (int ax = 0; ax < 3; ax++) {
        double dd = 0.0;
        int nbc = 0;
        const float own = ax == 0 ? n.x : (ax == 1 ? n.y : n.z);
        if (s.ex & (1u << (2 * ax))) {
          const V3 b = s.nb[2 * ax];
          dd += (double)((ax == 0 ? b.x : (ax == 1 ? b.y : b.z)) - own);
          nbc++;
        }
        if (s.ex & (1u << (2 * ax + 1))) {
          const V3 b = s.nb[2 * ax + 1];
          dd -= (double)((ax == 0 ? b.x : (ax == 1 ? b.y : b.z)) - own);
          nbc++;
        }
        if (nbc) div += dd / nbc;
      }
      q[20] += div;
    }
  }
  __shared__ double red[kWaves][kNQ];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < kNQ; a++) {
    if (!written_q<POT, XQ>(a)) continue;
    const double v = wave_sum(q[a]);
    if (lane == 0) red[w][a] = v;
  }
  __syncthreads();
  // every quantity k_finalize reads gets a partial; those this potential never writes are zero
  if (threadIdx.x < kNQ && (XQ || basic_q(threadIdx.x))) {
    double v = 0.0;
    if (written_q<POT, XQ>(threadIdx.x)) {
#pragma unroll
      for (int ww = 0; ww < kWaves; ww++) v += red[ww][threadIdx.x];
    }
    partials[(size_t)threadIdx.x * nmblk + Lb] = v;
  }
}

// Largest eigenvalue and eigenvector of a symmetric 3x3 matrix in FP64: trigonometric solution of the
// characteristic cubic (O. K. Smith, "Eigenvalues of a symmetric 3x3 matrix", CACM 4(4):168, 1961); the
// eigenvector is the largest cross product of two rows of A - lambda I.
__device__ __forceinline__ double eig_sym3_max(const double a[6], double v[3]) {
  // a = (xx, xy, xz, yy, yz, zz)
  const double a00 = a[0], a01 = a[1], a02 = a[2], a11 = a[3], a12 = a[4], a22 = a[5];
  const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  double lam;
  if (p1 == 0.0) {
    lam = a00;
    v[0] = 1; v[1] = 0; v[2] = 0;
    if (a11 > lam) { lam = a11; v[0] = 0; v[1] = 1; v[2] = 0; }
    if (a22 > lam) { lam = a22; v[0] = 0; v[1] = 0; v[2] = 1; }
    return lam;
  }
  const double q = (a00 + a11 + a22) / 3.0;
  const double p2 = (a00 - q) * (a00 - q) + (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + 2.0 * p1;
  const double p = sqrt(p2 / 6.0);
  const double b00 = (a00 - q) / p, b11 = (a11 - q) / p, b22 = (a22 - q) / p;
  const double b01 = a01 / p, b02 = a02 / p, b12 = a12 / p;
  double rr = 0.5 * (b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02));
  rr = fmin(1.0, fmax(-1.0, rr));
  const double phi = acos(rr) / 3.0;
  lam = q + 2.0 * p * cos(phi);
  // cross products of the rows of A - lambda I (scalars, so nothing is indexed at run time)
  const double r00 = a00 - lam, r11 = a11 - lam, r22 = a22 - lam;
  const double c00 = a01 * a12 - a02 * r11, c01 = a02 * a01 - r00 * a12, c02 = r00 * r11 - a01 * a01;  // r0 x r1
  const double c10 = a01 * r22 - a02 * a12, c11 = a02 * a02 - r00 * r22, c12 = r00 * a12 - a01 * a02;  // r0 x r2
  const double c20 = r11 * r22 - a12 * a12, c21 = a12 * a02 - a01 * r22, c22 = a01 * a12 - r11 * a02;  // r1 x r2
  const double n0 = c00 * c00 + c01 * c01 + c02 * c02;
  const double n1 = c10 * c10 + c11 * c11 + c12 * c12;
  const double n2 = c20 * c20 + c21 * c21 + c22 * c22;
  double x = c00, y = c01, z = c02, bn = n0;
  if (n1 > bn) { x = c10; y = c11; z = c12; bn = n1; }
  if (n2 > bn) { x = c20; y = c21; z = c22; bn = n2; }
  if (bn > 0.0) {
    const double inv = 1.0 / sqrt(bn);
    v[0] = x * inv; v[1] = y * inv; v[2] = z * inv;
  } else {
    v[0] = 0; v[1] = 0; v[2] = 1;
  }
  return lam;
}

// Sums the partials of each lattice in a fixed order (one workgroup per lattice), computes the observables and,
// when accumulate != 0, adds the sample to the moments.
// Partials come either from k_measure (ranges == nullptr: the lattice's range [mp0, mp1), quantities 0..kNQ-1) or
// from the measuring sweep pass (ranges[r]: the lattice's workgroups of the sweep work list, compact layout of kNQS
// basic quantities, xq = 0).
__global__ __launch_bounds__(kFinBlock) void k_finalize(const LatDev *__restrict__ lats,
                                                     const double *__restrict__ partials, int nmblk, int xq,
                                                     MomDev *__restrict__ mom, ObsDev *__restrict__ obs,
                                                     int accumulate, const int2 *__restrict__ ranges, int l0,
                                                     unsigned long long *__restrict__ sweep_base, unsigned advance) {
  const int r = l0 + (int)blockIdx.x;
  const LatDev &L = lats[r];
  // the moments are loaded here, ahead of the partials and the barrier, so that the serial tail of thread 0 does not
  // wait for them (this kernel ends a measured call)
  const MomDev mprev = mom[r];
  __shared__ double red[kFinWaves][kNQ];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double v[kNQ];
#pragma unroll
  for (int a = 0; a < kNQ; a++) v[a] = 0.0;
  if (ranges) {
    // four partials per lane in flight, added in increasing order (the order finalize_compact keeps); a lattice
    // with a radial order has its sums in rows 8-10, added to the sphere (7-9) or cylinder (10-12) sums
    const int2 rg = ranges[r];
    const bool rad = radial_flags(L.flags);
    const int ro = (L.flags & F_RSPH) ? 7 : 10;
#pragma unroll 1
    for (int p = rg.x + (int)threadIdx.x; p < rg.y; p += 4 * kFinBlock) {
      double x[4][kNQSR];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = p + u * kFinBlock;
#pragma unroll
        for (int a = 0; a < kNQSR; a++)
          x[u][a] = q < rg.y && (a < kNQS || rad) ? partials[(size_t)a * nmblk + q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (p + u * kFinBlock >= rg.y) continue;
#pragma unroll
        for (int a = 0; a < 7; a++) v[a] += x[u][a];
        v[21] += x[u][7];
        if (rad) {
          v[ro] += x[u][8];
          v[ro + 1] += x[u][9];
          v[ro + 2] += x[u][10];
        }
      }
    }
  } else {
    for (int p = L.mp0 + (int)threadIdx.x; p < L.mp1; p += kFinBlock) {
#pragma unroll
      for (int a = 0; a < kNQ; a++)
        if (xq || basic_q(a)) v[a] += partials[(size_t)a * nmblk + p];
    }
  }
  // quantities never written are left out of the reduction: the radial and pear ones without xq (the compact
  // layout's radial rows come with xq), and the pear ones of the compact layout
  auto used = [&](int a) { return (xq || basic_q(a)) && !(ranges && a >= 13 && a <= 20); };
#pragma unroll
  for (int a = 0; a < kNQ; a++) {
    if (!used(a)) continue;
    const double x = wave_sum(v[a]);
    if (lane == 0) red[w][a] = x;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tot[kNQ];
  for (int a = 0; a < kNQ; a++) {
    double x = 0.0;
    if (used(a))
      for (int ww = 0; ww < kFinWaves; ww++) x += red[ww][a];
    tot[a] = x;
  }
  const double N = tot[6];
  ObsDev o;
  o.nvalid = N;
  const double inv = N > 0.0 ? 1.0 / N : 0.0;
  o.E = tot[0] * inv;
  // Q = <(3 n n - I) / 2>
  o.Q[0] = (3.0 * tot[1] - N) * 0.5 * inv;
  o.Q[1] = 3.0 * tot[2] * 0.5 * inv;
  o.Q[2] = 3.0 * tot[3] * 0.5 * inv;
  o.Q[3] = (3.0 * tot[4] - N) * 0.5 * inv;
  o.Q[4] = 3.0 * tot[5] * 0.5 * inv;
  o.Q[5] = (3.0 * tot[21] - N) * 0.5 * inv;
  // parameter_order.cpp:70 sets Q_zz = -Q_xx - Q_yy (unit directors); the eigen-solve uses the same convention
  double qm[6] = {o.Q[0], o.Q[1], o.Q[2], o.Q[3], o.Q[4], -o.Q[0] - o.Q[3]};
  o.S = N > 0.0 ? eig_sym3_max(qm, o.dir) : 0.0;
  if (N <= 0.0) { o.dir[0] = 0; o.dir[1] = 0; o.dir[2] = 1; }
  o.srs = tot[9] > 0.0 ? tot[7] / tot[9] : 0.0;
  o.prs = tot[9] > 0.0 ? tot[8] / tot[9] : 0.0;
  o.src = tot[12] > 0.0 ? tot[10] / tot[12] : 0.0;
  o.prc = tot[12] > 0.0 ? tot[11] / tot[12] : 0.0;
  o.P = fabs(o.dir[0] * tot[13] + o.dir[1] * tot[14] + o.dir[2] * tot[15]) * inv;
  o.dth = tot[19] > 0.0 ? 1.5 * sqrt(tot[16] * tot[16] + tot[17] * tot[17] + tot[18] * tot[18]) / tot[19] : 0.0;
  o.divn = tot[20] * inv;
  obs[r] = o;
  if (accumulate) {
    MomDev m = mprev;
    m.n += 1.0;
    m.S += o.S;
    m.S2 += o.S * o.S;
    m.E += o.E;
    m.E2 += o.E * o.E;
    m.srs += o.srs;
    m.prs += o.prs;
    m.src += o.src;
    m.prc += o.prc;
    m.P += o.P;
    m.dth += o.dth;
    m.divn += o.divn;
    mom[r] = m;
  }
  // the last finalize of a call advances the sweep base itself (instead of a k_advance launch after it)
  if (advance && blockIdx.x == 0) *sweep_base += advance;
}

__global__ void k_advance(unsigned long long *sweep_base, unsigned n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *sweep_base += n;
}

__global__ void k_zero_moments(MomDev *mom, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) mom[r] = MomDev{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
}

// Local order parameter of every site of one lattice (lattice_order_parameter, parameter_order.cpp:75-127):
// Q over the site and its non-empty neighbours, largest eigenvalue in FP64; NaN on empty sites. Output in the
// AoS site order idx = i + nx (j + ny k).
template <bool MASK>
__global__ __launch_bounds__(kBlock) void k_local_order(const LatDev *__restrict__ lats, int r,
                                                        const float *__restrict__ X, const float *__restrict__ Y,
                                                        const float *__restrict__ Z, const uint8_t *__restrict__ PT,
                                                        float *__restrict__ out) {
  const LatDev &L = lats[r];
  const int g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= 2 * L.nc) return;
  const int colour = g >= L.nc ? 1 : 0;
  const int t = g - colour * L.nc;
  Site s;
  uint32_t self;
  const uint32_t row = fdiv((uint32_t)t, L.mxh, L.sxh);
  const int m = t - (int)row * L.nxh;
  const int k = (int)fdiv(row, L.my, L.sy);
  const int j = (int)row - k * L.ny;
  const int i = 2 * m + ((j + k + colour) & 1);
  if (i >= L.nx) return;
  const long long lin = (long long)i + (long long)L.nx * (j + (long long)L.ny * k);
  if (!gather<MASK, false>(L, colour, t, X, Y, Z, PT, nullptr, nullptr, nullptr, s, self)) {
    out[lin] = __int_as_float(0x7fc00000);
    return;
  }
  double a[6] = {0, 0, 0, 0, 0, 0};
  int pts = 0;
  auto add = [&](V3 v) {
    a[0] += 3.0 * (double)(v.x * v.x) - 1.0;
    a[1] += 3.0 * (double)(v.x * v.y);
    a[2] += 3.0 * (double)(v.x * v.z);
    a[3] += 3.0 * (double)(v.y * v.y) - 1.0;
    a[4] += 3.0 * (double)(v.y * v.z);
    a[5] += 3.0 * (double)(v.z * v.z) - 1.0;
    pts++;
  };
  add(s.n);
  for (int d = 0; d < 6; d++)
    if (s.ex & (1u << d)) add(s.nb[d]);
  for (int x = 0; x < 6; x++) a[x] /= 2.0 * pts;
  double v[3];
  out[lin] = (float)eig_sym3_max(a, v);
}

// Energies of given sites with their current director and with a trial director, by the sweep's energy function.
template <int POT, bool MASK, bool SURF, bool FIELD>
__global__ void k_site_energies(const LatDev *__restrict__ lats, const CtrlDev *__restrict__ ctrl, int r,
                                const int *__restrict__ sites, const float *__restrict__ trial, int n,
                                const float *__restrict__ X, const float *__restrict__ Y,
                                const float *__restrict__ Z, const uint8_t *__restrict__ PT,
                                const float *__restrict__ SX, const float *__restrict__ SY,
                                const float *__restrict__ SZ, const SurfDev *__restrict__ surf,
                                float *__restrict__ e_cur, float *__restrict__ e_trial) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const LatDev &L = lats[r];
  const int lin = sites[g];
  const int i = lin % L.nx, j = (lin / L.nx) % L.ny, k = lin / (L.nx * L.ny);
  const int colour = (i + j + k) & 1;
  const int t = (i >> 1) + L.nxh * (j + L.ny * k);
  Site s;
  uint32_t self;
  if (!gather<MASK, SURF>(L, colour, t, X, Y, Z, PT, SX, SY, SZ, s, self)) {
    e_cur[g] = 0.0f;
    e_trial[g] = 0.0f;
    return;
  }
  const float elecE = ctrl[r].elecE;
  e_cur[g] = site_energy<POT, SURF, FIELD>(L, s, s.n, surf, elecE);
  const V3 tv{trial[3 * g], trial[3 * g + 1], trial[3 * g + 2]};
  e_trial[g] = site_energy<POT, SURF, FIELD>(L, s, tv, surf, elecE);
}

// Plain copy (16 bytes per lane per iteration, grid-stride): the STREAM-style kernel used to confirm the HBM peak
// and to calibrate the FETCH_SIZE / WRITE_SIZE counters against a known number of bytes.
__global__ __launch_bounds__(kBlock) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x; g < n; g += stride) dst[g] = src[g];
}

__global__ void k_philox(const uint32_t *__restrict__ ctr, const uint32_t *__restrict__ key, int n,
                         uint32_t *__restrict__ out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const mcl::Philox4 R =
      mcl::philox4x32_10(ctr[4 * g], ctr[4 * g + 1], ctr[4 * g + 2], ctr[4 * g + 3], key[2 * g], key[2 * g + 1]);
  for (int x = 0; x < 4; x++) out[4 * g + x] = R.v[x];
}

// ---------------------------------------------------------------------------------------------------------------
// Host side.

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define MCL_HIP(call)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return fail(e_ == hipErrorOutOfMemory ? MCL_ERR_NOMEM : MCL_ERR_DEVICE,                         \
                  std::string(#call) + ": " + hipGetErrorString(e_));                                 \
  } while (0)

struct LatHost {
  int nx, ny, nz, nxh, nc;
  long long off;
  int nvalid;
  bool mask, surf;
  int radial;
  // LL context (Ctx::zero_empty): the empty sites (linear index) are zero vectors on the device; the directors the
  // caller gave them are kept here and returned by mcl_get_directors
  std::vector<uint32_t> empty;
  std::vector<float> empty_ni;
};

// Lattice groups. A context whose lattices together hold more than about the 256 MB Infinity Cache streams every
// colour pass from HBM: a batch of 64^3 lattices runs at 157 site-updates/ns with 252 MB of directors and at 109 with
// 302 MB. mcl_sweeps therefore runs such a context group by group, each group a run of consecutive lattices of at most
// group_bytes, through all the sweeps of the call before the next group starts. Every lattice's trajectory depends
// only on its own sites and sweep numbers, so the order of the groups changes nothing in the results.
// A group: lattices [l0, l1), their workgroups [b0, b1) of the sweep work list and [m0, m1) of the measurement list;
// chunk: the tiles per XCD of an XCD-ordered launch over the group; share: its part of the context's positions.
struct LatGroup {
  int l0 = 0, l1 = 0, b0 = 0, b1 = 0, m0 = 0, m1 = 0, chunk = 0;
  double share = 1.0;
};
// Default largest group: directors and point types of both colours.
constexpr long long kGroupBytes = 192ll << 20;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nlat = 0;
  int pot = 0;
  bool mask = false, surf = false, field = false, flip = false;
  bool free = false;  // some lattice has a free (non-periodic) axis
  bool row = false;   // every lattice is row-aligned: nx a multiple of 128 (gather_row)
  bool zero_empty = false;  // LL context with masks: empty sites are stored as zero vectors (see gather's NBPT)
  bool resident_ok = false;  // every lattice fits the LDS of one workgroup (k_sweep_resident)
  bool resident_on = true;   // plain sweeps of such a context run resident (mcl_set_resident)
  bool resident_one = false;  // ... with at most kResBlock positions per colour and no flip trial (ONE)
  bool xq = false;  // radial or pear observables are measured
  std::vector<LatHost> lh;
  std::vector<LatDev> ld;
  long long pool = 0;  // sites in the pooled colour blocks
  float *X = nullptr, *Y = nullptr, *Z = nullptr;
  uint8_t *PT = nullptr;
  float *SX = nullptr, *SY = nullptr, *SZ = nullptr;
  SurfDev *surf_d = nullptr;
  LatDev *lat_d = nullptr;
  CtrlDev *ctrl_d = nullptr;
  int2 *blk_d = nullptr;
  unsigned int *cnt = nullptr;  // striped acceptance counters [lattice][sweep % 3][kCnt]
  unsigned long long sweeps_done = 0;            // host copy of the device sweep base
  unsigned long long *sweep_base = nullptr;      // device sweep base (advanced at the end of every mcl_sweeps)
  int nblk = 0;  // workgroups of the sweep work list
  int bs = 256, xcd = 1;  // sweep launch shape: threads per workgroup, XCD-contiguous tile order
  int graph_block = 0;    // > 0: sweeps replayed from captured graphs of about this many sweeps
  std::vector<std::pair<std::pair<int, int>, hipGraphExec_t>> graphs;  // ((sweeps, measure_every), graph)
  int4 *mblk_d = nullptr;
  int nmblk = 0;
  double *partials = nullptr;
  // measurement folded into the colour-1 sweep pass (every context but pear, GHRL with empty sites, lattices one
  // site wide in x, and lattices measuring both radial orders); fr: it carries the radial order (kNQSR partials)
  bool fused = false;
  bool fr = false;
  bool has_empty = false;  // some lattice has an empty site (masked GHRL / pear sweeps then check neighbours)
  bool split = false;      // one position per sweep thread (SPLIT): a small, not row-aligned context without flips
  double *spart = nullptr;  // kNQS (kNQSR with fr) partials per workgroup of the sweep work list
  size_t spart_cap = 0;     // workgroups spart holds
  int2 *sr_d = nullptr;     // per lattice: its workgroups [first, last) in the sweep work list
  MomDev *mom_d = nullptr;
  ObsDev *obs_d = nullptr;
  float *scratch_f = nullptr;  // Nt floats of the largest lattice (local order output)
  long long max_nt = 0;
  // kernel timing
  int timing_stride = 0;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  // (first event of the pair, colour passes over the whole context in the run: a pass over a lattice group counts
  // as its share)
  std::vector<std::pair<size_t, double>> ev_runs;
  long long group_bytes = kGroupBytes;  // <= 0: one group
  std::vector<LatGroup> groups;         // the groups mcl_sweeps runs in turn (one when the context fits)
  LatGroup all;                         // the whole context
};

using SweepFn = void (*)(dim3, dim3, hipStream_t, const Ctx &, const LatGroup &, int, unsigned, bool, int);
using MeasFn = void (*)(dim3, dim3, hipStream_t, const Ctx &, int);

template <int POT, bool MASK, bool SURF, bool FIELD, bool FLIP, bool FREE, bool MEAS, bool ROW, bool XR = false,
          bool SPLIT = false>
void launch_sweep_k(dim3 g, dim3 b, hipStream_t st, const Ctx &c, const LatGroup &gr, int colour, unsigned k,
                    int fin) {
  if constexpr (!FLIP && !ROW && !SPLIT) {
    if (c.split) {
      launch_sweep_k<POT, MASK, SURF, FIELD, FLIP, FREE, MEAS, ROW, XR, true>(g, b, st, c, gr, colour, k, fin);
      return;
    }
  }
  const int nfin = c.fused ? gr.l1 - gr.l0 : 0;
  // the measuring pass of an unmasked LL context reduces its partials through dynamic LDS (at most 32 KiB)
  constexpr bool LDSRED = MEAS && !MASK && !XR && POT == MCL_POT_LL;
  const unsigned shm = LDSRED ? (unsigned)(kNQS * sizeof(double)) * (b.x <= 512 ? b.x : b.x / 2) : 0u;
  if constexpr (MASK && POT != MCL_POT_LL) {
    if (!c.has_empty) {
      hipLaunchKernelGGL((k_sweep<POT, MASK, SURF, FIELD, FLIP, FREE, MEAS, ROW, false, SPLIT, XR>), g, b, shm, st,
                         c.lat_d, c.ctrl_d, c.cnt, c.blk_d, c.nblk, gr.b0, gr.b1, gr.chunk, c.xcd, colour,
                         c.sweep_base, k, c.X, c.Y, c.Z, c.PT, c.SX, c.SY, c.SZ, c.surf_d, c.spart, fin, gr.l0, nfin,
                         c.sr_d, c.mom_d, c.obs_d);
      return;
    }
  }
  hipLaunchKernelGGL((k_sweep<POT, MASK, SURF, FIELD, FLIP, FREE, MEAS, ROW, true, SPLIT, XR>), g, b, shm, st, c.lat_d,
                     c.ctrl_d, c.cnt, c.blk_d, c.nblk, gr.b0, gr.b1, gr.chunk, c.xcd, colour, c.sweep_base, k, c.X,
                     c.Y, c.Z, c.PT, c.SX, c.SY, c.SZ, c.surf_d, c.spart, fin, gr.l0, nfin, c.sr_d, c.mom_d, c.obs_d);
}

// meas (the measurement folded into the colour-1 pass) is never set for pear contexts: they always measure the pear
// observables, which keep the separate measurement kernels, so no pear measuring instance is compiled. A context
// that measures a radial order takes the generic gather for its measuring passes (the same trials as the row path).
template <int POT, bool MASK, bool SURF, bool FIELD, bool FLIP, bool FREE>
void launch_sweep(dim3 g, dim3 b, hipStream_t st, const Ctx &c, const LatGroup &gr, int colour, unsigned k, bool meas,
                  int fin) {
  if constexpr (POT != MCL_POT_PEAR) {
    if (meas) {
      if (c.fr)
        launch_sweep_k<POT, MASK, SURF, FIELD, FLIP, FREE, true, false, true>(g, b, st, c, gr, colour, k, fin);
      else if (c.row)
        launch_sweep_k<POT, MASK, SURF, FIELD, FLIP, FREE, true, true>(g, b, st, c, gr, colour, k, fin);
      else
        launch_sweep_k<POT, MASK, SURF, FIELD, FLIP, FREE, true, false>(g, b, st, c, gr, colour, k, fin);
      return;
    }
  }
  if (c.row)
    launch_sweep_k<POT, MASK, SURF, FIELD, FLIP, FREE, false, true>(g, b, st, c, gr, colour, k, fin);
  else
    launch_sweep_k<POT, MASK, SURF, FIELD, FLIP, FREE, false, false>(g, b, st, c, gr, colour, k, fin);
}

// Grid of a sweep launch over a lattice group: its part of the work list, plus one workgroup per lattice of the
// group for the deferred finalize of a measurement folded into the previous sweep.
dim3 sweep_grid(const Ctx *c, const LatGroup &gr) {
  return dim3(c->xcd ? 8 * gr.chunk : (gr.b1 - gr.b0) + (c->fused ? gr.l1 - gr.l0 : 0));
}
using ResFn = void (*)(hipStream_t, const Ctx &, unsigned, int);
template <int POT, bool MASK, bool SURF, bool FIELD, bool FLIP, bool FREE, bool ONE>
void launch_resident_k(hipStream_t st, const Ctx &c, unsigned k0, int nsw) {
  if constexpr (MASK && POT != MCL_POT_LL) {
    if (!c.has_empty) {
      hipLaunchKernelGGL((k_sweep_resident<POT, MASK, SURF, FIELD, FLIP, FREE, ONE, false>), dim3(c.nlat),
                         dim3(kResBlock), 0, st, c.lat_d, c.ctrl_d, c.cnt, c.sweep_base, k0, nsw, c.X, c.Y, c.Z,
                         c.PT, c.SX, c.SY, c.SZ, c.surf_d);
      return;
    }
  }
  hipLaunchKernelGGL((k_sweep_resident<POT, MASK, SURF, FIELD, FLIP, FREE, ONE>), dim3(c.nlat), dim3(kResBlock), 0,
                     st, c.lat_d, c.ctrl_d, c.cnt, c.sweep_base, k0, nsw, c.X, c.Y, c.Z, c.PT, c.SX, c.SY, c.SZ,
                     c.surf_d);
}
template <int POT, bool MASK, bool SURF, bool FIELD, bool FLIP, bool FREE>
void launch_resident(hipStream_t st, const Ctx &c, unsigned k0, int nsw) {
  if constexpr (!FLIP) {
    if (c.resident_one) {
      launch_resident_k<POT, MASK, SURF, FIELD, FLIP, FREE, true>(st, c, k0, nsw);
      return;
    }
  }
  launch_resident_k<POT, MASK, SURF, FIELD, FLIP, FREE, false>(st, c, k0, nsw);
}
template <int POT, bool MASK, bool SURF, bool FREE>
ResFn pick_resident_field(bool field, bool flip) {
  if (field)
    return flip ? launch_resident<POT, MASK, SURF, true, true, FREE> : launch_resident<POT, MASK, SURF, true, false, FREE>;
  return flip ? launch_resident<POT, MASK, SURF, false, true, FREE> : launch_resident<POT, MASK, SURF, false, false, FREE>;
}
// the same instance choice as pick_sweep
template <int POT>
ResFn pick_resident_geom(bool mask, bool surf, bool free, bool field, bool flip) {
  if (surf) return pick_resident_field<POT, true, true, true>(field, flip);
  if (mask) return pick_resident_field<POT, true, false, true>(field, flip);
  if (free) return pick_resident_field<POT, false, false, true>(field, flip);
  return pick_resident_field<POT, false, false, false>(field, flip);
}
ResFn pick_resident(const Ctx &c) {
  switch (c.pot) {
    case MCL_POT_GHRL: return pick_resident_geom<MCL_POT_GHRL>(c.mask, c.surf, c.free, c.field, c.flip);
    case MCL_POT_PEAR: return pick_resident_geom<MCL_POT_PEAR>(c.mask, c.surf, c.free, c.field, c.flip);
    default: return pick_resident_geom<MCL_POT_LL>(c.mask, c.surf, c.free, c.field, c.flip);
  }
}

template <int POT, bool MASK, bool SURF, bool FIELD, bool FREE>
void launch_measure(dim3 g, dim3 b, hipStream_t st, const Ctx &c, int m0) {
  if (c.xq)
    hipLaunchKernelGGL((k_measure<POT, MASK, SURF, FIELD, true, FREE>), g, b, 0, st, c.lat_d, c.ctrl_d, c.mblk_d,
                       c.nmblk, m0, c.X, c.Y, c.Z, c.PT, c.SX, c.SY, c.SZ, c.surf_d, c.partials);
  else
    hipLaunchKernelGGL((k_measure<POT, MASK, SURF, FIELD, false, FREE>), g, b, 0, st, c.lat_d, c.ctrl_d, c.mblk_d,
                       c.nmblk, m0, c.X, c.Y, c.Z, c.PT, c.SX, c.SY, c.SZ, c.surf_d, c.partials);
}

template <int POT, bool MASK, bool SURF, bool FIELD, bool FREE>
SweepFn pick_sweep_flip(bool flip) {
  return flip ? launch_sweep<POT, MASK, SURF, FIELD, true, FREE> : launch_sweep<POT, MASK, SURF, FIELD, false, FREE>;
}
template <int POT, bool MASK, bool SURF, bool FREE>
SweepFn pick_sweep_field(bool field, bool flip) {
  return field ? pick_sweep_flip<POT, MASK, SURF, true, FREE>(flip)
               : pick_sweep_flip<POT, MASK, SURF, false, FREE>(flip);
}
template <int POT>
SweepFn pick_sweep_geom(bool mask, bool surf, bool free, bool field, bool flip) {
  if (surf) return pick_sweep_field<POT, true, true, true>(field, flip);
  if (mask) return pick_sweep_field<POT, true, false, true>(field, flip);
  if (free) return pick_sweep_field<POT, false, false, true>(field, flip);
  return pick_sweep_field<POT, false, false, false>(field, flip);
}
SweepFn pick_sweep(const Ctx &c) {
  switch (c.pot) {
    case MCL_POT_GHRL: return pick_sweep_geom<MCL_POT_GHRL>(c.mask, c.surf, c.free, c.field, c.flip);
    case MCL_POT_PEAR: return pick_sweep_geom<MCL_POT_PEAR>(c.mask, c.surf, c.free, c.field, c.flip);
    default: return pick_sweep_geom<MCL_POT_LL>(c.mask, c.surf, c.free, c.field, c.flip);
  }
}

template <int POT, bool MASK, bool SURF, bool FREE>
MeasFn pick_meas_field(bool field) {
  return field ? launch_measure<POT, MASK, SURF, true, FREE> : launch_measure<POT, MASK, SURF, false, FREE>;
}
template <int POT>
MeasFn pick_meas_geom(bool mask, bool surf, bool free, bool field) {
  if (surf) return pick_meas_field<POT, true, true, true>(field);
  if (mask) return pick_meas_field<POT, true, false, true>(field);
  if (free) return pick_meas_field<POT, false, false, true>(field);
  return pick_meas_field<POT, false, false, false>(field);
}
MeasFn pick_meas(const Ctx &c) {
  switch (c.pot) {
    case MCL_POT_GHRL: return pick_meas_geom<MCL_POT_GHRL>(c.mask, c.surf, c.free, c.field);
    case MCL_POT_PEAR: return pick_meas_geom<MCL_POT_PEAR>(c.mask, c.surf, c.free, c.field);
    default: return pick_meas_geom<MCL_POT_LL>(c.mask, c.surf, c.free, c.field);
  }
}

Ctx *as_ctx(void *p) { return static_cast<Ctx *>(p); }

int check_lat(Ctx *c, int lat) {
  if (!c) return fail(MCL_ERR_ARG, "null context");
  if (lat < 0 || lat >= c->nlat) return fail(MCL_ERR_ARG, "lattice index out of range");
  return MCL_OK;
}

// AoS host directors <-> pooled colour-split SoA of lattice r.
void aos_to_soa(const LatHost &h, const float *ni, std::vector<float> &x, std::vector<float> &y,
                std::vector<float> &z) {
  x.assign(2 * (size_t)h.nc, 0.0f);
  y.assign(2 * (size_t)h.nc, 0.0f);
  z.assign(2 * (size_t)h.nc, 1.0f);
  for (int k = 0; k < h.nz; k++)
    for (int j = 0; j < h.ny; j++)
      for (int i = 0; i < h.nx; i++) {
        const size_t lin = (size_t)i + (size_t)h.nx * (j + (size_t)h.ny * k);
        const int colour = (i + j + k) & 1;
        const size_t t = (size_t)colour * h.nc + (i >> 1) + (size_t)h.nxh * (j + (size_t)h.ny * k);
        x[t] = ni[3 * lin];
        y[t] = ni[3 * lin + 1];
        z[t] = ni[3 * lin + 2];
      }
}

int upload_directors(Ctx *c, int r, const float *ni) {
  LatHost &h = c->lh[r];
  std::vector<float> x, y, z;
  aos_to_soa(h, ni, x, y, z);
  if (c->zero_empty && !h.empty.empty()) {
    // keep the caller's directors of the empty sites, store zero vectors
    h.empty_ni.resize(3 * h.empty.size());
    for (size_t e = 0; e < h.empty.size(); e++) {
      const uint32_t lin = h.empty[e];
      for (int a = 0; a < 3; a++) h.empty_ni[3 * e + a] = ni[3 * (size_t)lin + a];
      const int i = (int)(lin % (uint32_t)h.nx), j = (int)((lin / (uint32_t)h.nx) % (uint32_t)h.ny);
      const int k = (int)(lin / ((uint32_t)h.nx * (uint32_t)h.ny));
      const size_t t = (size_t)((i + j + k) & 1) * h.nc + (i >> 1) + (size_t)h.nxh * (j + (size_t)h.ny * k);
      x[t] = y[t] = z[t] = 0.0f;
    }
  }
  const size_t bytes = sizeof(float) * 2 * (size_t)h.nc;
  MCL_HIP(hipMemcpyAsync(c->X + h.off, x.data(), bytes, hipMemcpyHostToDevice, c->stream));
  MCL_HIP(hipMemcpyAsync(c->Y + h.off, y.data(), bytes, hipMemcpyHostToDevice, c->stream));
  MCL_HIP(hipMemcpyAsync(c->Z + h.off, z.data(), bytes, hipMemcpyHostToDevice, c->stream));
  MCL_HIP(hipStreamSynchronize(c->stream));
  return MCL_OK;
}

void destroy_ctx(Ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  void *bufs[] = {c->X, c->Y, c->Z, c->PT, c->SX, c->SY, c->SZ, c->surf_d, c->lat_d, c->ctrl_d,
                  c->blk_d, c->cnt, c->sweep_base, c->mblk_d, c->partials, c->mom_d, c->obs_d, c->scratch_f,
                  c->spart, c->sr_d};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Staging of mcl_site_energies_noalloc: MCL_SCRATCH_SITES site indices, trial directors and the two energies (24 B per
// site, 1.5 MiB) in one file-scope device array. Each GPU the library runs on holds its own copy until the process
// ends; every call on a GPU uses that GPU's copy, so calls that overlap on one GPU race. On the host the name is only
// a stand-in for the array: its device address comes from hipGetSymbolAddress, looked up once per device (with that
// device current) when the first context is created there, and only that address is handed to copies and kernels.
constexpr size_t kScratchBytes = (size_t)MCL_SCRATCH_SITES * (sizeof(int) + 5 * sizeof(float));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_dev;  // device address of `scratch`, by device index; nullptr until looked up

// Looks up the current device's address of `scratch` (device index `device`) unless it is known.
static int scratch_lookup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize((size_t)device + 1, nullptr);
  if (g_scratch_dev[device]) return MCL_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(MCL_ERR_DEVICE, std::string("scratch address: ") + hipGetErrorString(e));
  g_scratch_dev[device] = ptr;
  return MCL_OK;
}

static void *scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch_dev.size() ? g_scratch_dev[device] : nullptr;
}

template <typename T>
int dev_alloc(T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess) return fail(MCL_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  *p = static_cast<T *>(q);
  return MCL_OK;
}

// Sweep work list for workgroups of bs threads: per lattice ceil(nc / bs) workgroups of bs consecutive colour
// positions. Drops the captured graphs, which hold the old launch shape.
int build_worklist(Ctx *c, int bs, int xcd) {
  // a workgroup of the sweep covers bs colour-block positions, or 2 bs when a thread updates a pair of sites
  const int span = c->flip || c->split ? bs : 2 * bs;
  std::vector<int2> blk, sr(c->nlat);
  for (int r = 0; r < c->nlat; r++) {
    sr[r].x = (int)blk.size();
    for (int t0 = 0; t0 < c->ld[r].nc; t0 += span) blk.push_back(make_int2(r, t0));
    sr[r].y = (int)blk.size();
  }
  if (c->blk_d) {
    MCL_HIP(hipStreamSynchronize(c->stream));
    MCL_HIP(hipFree(c->blk_d));
    c->blk_d = nullptr;
  }
  int st;
  if ((st = dev_alloc(&c->blk_d, blk.size()))) return st;
  MCL_HIP(hipMemcpy(c->blk_d, blk.data(), sizeof(int2) * blk.size(), hipMemcpyHostToDevice));
  if (c->fused) {
    if (blk.size() > c->spart_cap) {
      MCL_HIP(hipStreamSynchronize(c->stream));
      if (c->spart) MCL_HIP(hipFree(c->spart));
      c->spart = nullptr;
      c->spart_cap = 0;
      if ((st = dev_alloc(&c->spart, (size_t)(c->fr ? kNQSR : kNQS) * blk.size()))) return st;
      c->spart_cap = blk.size();
    }
    if (!c->sr_d && (st = dev_alloc(&c->sr_d, (size_t)c->nlat))) return st;
    MCL_HIP(hipMemcpy(c->sr_d, sr.data(), sizeof(int2) * sr.size(), hipMemcpyHostToDevice));
  }
  c->nblk = (int)blk.size();
  c->bs = bs;
  c->xcd = xcd;
  // lattice groups: consecutive lattices, about equal in bytes, as few as keep each within group_bytes
  long long all_pos = 0;
  for (int r = 0; r < c->nlat; r++) all_pos += c->ld[r].nc;
  auto make_group = [&](int l0, int l1) {
    LatGroup g;
    g.l0 = l0;
    g.l1 = l1;
    g.b0 = sr[l0].x;
    g.b1 = sr[l1 - 1].y;
    g.m0 = c->ld[l0].mp0;
    g.m1 = c->ld[l1 - 1].mp1;
    g.chunk = (g.b1 - g.b0 + (c->fused ? l1 - l0 : 0) + 7) / 8;
    long long pos = 0;
    for (int r = l0; r < l1; r++) pos += c->ld[r].nc;
    g.share = (double)pos / (double)all_pos;
    return g;
  };
  c->all = make_group(0, c->nlat);
  c->groups.clear();
  std::vector<long long> bytes(c->nlat);
  long long total = 0;
  for (int r = 0; r < c->nlat; r++) total += bytes[r] = 2ll * c->ld[r].nc * (12 + (c->mask ? 1 : 0));
  const long long ng = c->group_bytes > 0 ? std::min<long long>(c->nlat, (total + c->group_bytes - 1) / c->group_bytes) : 1;
  if (ng <= 1) {
    c->groups.push_back(c->all);
  } else {
    const double target = (double)total / (double)ng;
    long long cum = 0;
    int l0 = 0;
    for (int r = 0; r < c->nlat; r++) {
      cum += bytes[r];
      // close the group at the lattice whose middle passes the group's share of the bytes
      if ((long long)c->groups.size() < ng - 1 && (double)(cum - bytes[r] / 2) >= target * (double)(c->groups.size() + 1)) {
        c->groups.push_back(make_group(l0, r + 1));
        l0 = r + 1;
      }
    }
    if (l0 < c->nlat) c->groups.push_back(make_group(l0, c->nlat));
  }
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  return MCL_OK;
}

int create_ctx(const mcl_lattice_desc *lat, int nlat, Ctx *c) {
  if (!lat || nlat < 1) return fail(MCL_ERR_ARG, "no lattices");
  MCL_HIP(hipGetDevice(&c->device));
  if (int st = scratch_lookup(c->device)) return st;
  MCL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->nlat = nlat;
  c->pot = lat[0].model.potential;
  c->flip = lat[0].model.flip_move != 0;
  if (c->pot < MCL_POT_LL || c->pot > MCL_POT_PEAR) return fail(MCL_ERR_ARG, "unknown potential");
  std::vector<SurfDev> surf_h;
  c->lh.resize(nlat);
  c->ld.resize(nlat);
  long long off = 0;
  for (int r = 0; r < nlat; r++) {
    const mcl_lattice_desc &d = lat[r];
    if (d.model.potential != c->pot || (d.model.flip_move != 0) != c->flip)
      return fail(MCL_ERR_ARG, "potential and flip_move must be the same for every lattice of a context");
    if (d.nx < 1 || d.ny < 1 || d.nz < 1) return fail(MCL_ERR_ARG, "grid sizes must be >= 1");
    if ((long long)d.nx * d.ny * d.nz >= (1ll << 31)) return fail(MCL_ERR_ARG, "lattice too large (>= 2^31 sites)");
    const int n3[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; a++)
      if (d.periodic[a] && (n3[a] % 2))
        return fail(MCL_ERR_ARG, "a periodic axis needs an even number of sites (checkerboard update)");
    if (!d.ni) return fail(MCL_ERR_ARG, "null director array");
    if (d.nsurf < 0 || d.nsurf > 253) return fail(MCL_ERR_ARG, "nsurf must be in [0, 253]");
    if (d.nsurf > 0 && (!d.surf || !d.ns)) return fail(MCL_ERR_ARG, "surfaces need the surface table and normals");
    LatHost &h = c->lh[r];
    h.nx = d.nx;
    h.ny = d.ny;
    h.nz = d.nz;
    h.nxh = (d.nx + 1) / 2;
    h.nc = h.nxh * d.ny * d.nz;
    h.off = off;
    off += 2ll * h.nc;
    const long long Nt = (long long)d.nx * d.ny * d.nz;
    c->max_nt = std::max(c->max_nt, Nt);
    h.nvalid = 0;
    h.mask = false;
    h.surf = false;
    if (d.pt) {
      for (long long s = 0; s < Nt; s++) {
        const int p = d.pt[s];
        if (p < 0 || p > 255) return fail(MCL_ERR_ARG, "point type out of range [0, 255]");
        if (p >= 2 && p - 2 >= d.nsurf) return fail(MCL_ERR_ARG, "point type names a surface without anchoring");
        h.nvalid += p ? 1 : 0;
        if (p != 1) h.mask = true;
        if (p >= 2) h.surf = true;
      }
    } else {
      h.nvalid = (int)Nt;
    }
    h.radial = d.radial;
    LatDev &L = c->ld[r];
    std::memset(&L, 0, sizeof(L));
    L.nx = d.nx;
    L.ny = d.ny;
    L.nz = d.nz;
    L.nxh = h.nxh;
    L.nc = h.nc;
    const FastDiv fx = make_fastdiv((uint32_t)h.nxh), fy = make_fastdiv((uint32_t)d.ny);
    L.mxh = fx.m;
    L.sxh = fx.sh;
    L.my = fy.m;
    L.sy = fy.sh;
    L.plane = (uint32_t)h.nxh * (uint32_t)d.ny;
    L.ywrap = (uint32_t)(d.ny - 1) * (uint32_t)h.nxh;
    L.zwrap = (uint32_t)(d.nz - 1) * L.plane;
    L.flags = (d.periodic[0] ? F_PX : 0) | (d.periodic[1] ? F_PY : 0) | (d.periodic[2] ? F_PZ : 0) |
              (h.mask ? F_MASK : 0) | (h.surf ? F_SURF : 0) | ((d.radial & 1) ? F_RSPH : 0) |
              ((d.radial & 2) ? F_RCYL : 0);
    L.nvalid = h.nvalid;
    L.off = (uint32_t)h.off;
    L.key0 = (uint32_t)d.seed;
    L.key1 = (uint32_t)(d.seed >> 32) ^ 0x5851F42Du;  // the stream depends on the seed only, not on the position
    L.surf0 = (int)surf_h.size();
    const mcl_model &M = d.model;
    L.A = M.A;
    L.B1 = M.B1;
    L.C = M.C;
    L.gl = M.ghrl_lambda;
    L.gm = M.ghrl_mu;
    L.gn = M.ghrl_nu;
    L.gr = M.ghrl_rho;
    L.gs = M.ghrl_sigma;
    L.ex = M.elec_dir[0];
    L.ey = M.elec_dir[1];
    L.ez = M.elec_dir[2];
    L.elecA = M.elecA;
    L.elecP = M.elecP;
    L.cx = (float)(d.nx / 2);
    L.cy = (float)(d.ny / 2);
    L.cz = (float)(d.nz / 2);
    if (M.elecA != 0.0f || M.elecP != 0.0f) c->field = true;
    if (!d.periodic[0] || !d.periodic[1] || !d.periodic[2]) c->free = true;
    if (d.radial != 0 || c->pot == MCL_POT_PEAR) c->xq = true;
    c->mask |= h.mask;
    c->surf |= h.surf;
    for (int s = 0; s < d.nsurf; s++) {
      const mcl_surface &S = d.surf[s];
      SurfDev sd{0, S.W, S.axis[0], S.axis[1], S.axis[2]};
      switch (S.type) {
        case MCL_ANCH_HOMEOTROPIC:
        case MCL_ANCH_RP:
        case MCL_ANCH_RP_TANGENT: sd.kind = 0; break;
        case MCL_ANCH_FG: sd.kind = 1; break;
        case MCL_ANCH_HOMEOTROPIC_GHRL: sd.kind = 2; break;
        case MCL_ANCH_RP_GHRL: sd.kind = 3; break;
        case MCL_ANCH_FG_GHRL: sd.kind = 4; break;
        default: return fail(MCL_ERR_ARG, "unknown anchoring type");
      }
      surf_h.push_back(sd);
    }
    if (d.nsurf > 0) {
      const SurfDev &f = surf_h[L.surf0];
      bool same = true;
      for (int s = 1; s < d.nsurf; s++) {
        const SurfDev &g = surf_h[L.surf0 + s];
        same = same && g.kind == f.kind && g.W == f.W && g.ax == f.ax && g.ay == f.ay && g.az == f.az;
      }
      if (same) {
        L.flags |= F_SURF1;
        L.s1 = f;
      }
    }
  }
  c->pool = off;
  // sweep work list: per lattice ceil(nc / kBlock) workgroups of kBlock consecutive colour positions
  if (off >= (1ll << 30)) return fail(MCL_ERR_ARG, "too many sites in one context (>= 2^30)");
  // measurement work list: per lattice and colour, workgroups of kBlock * kMeasPerThread positions. A context that
  // would get fewer than 2048 such workgroups (small lattices, few of them) takes spans down to kBlock instead, so
  // that a colour gets at least 64 workgroups where the lattice allows: its measurement would otherwise be a few
  // workgroups whose threads visit 8 sites each in turn (pear 20^3: 28.8 -> 22.9 us per measured sweep). A large
  // batch keeps the long spans, whose fewer partials cost less to finalize.
  std::vector<int4> mblk;
  long long nlong = 0;
  for (int r = 0; r < nlat; r++) nlong += 2ll * ((c->ld[r].nc + kBlock * kMeasPerThread - 1) / (kBlock * kMeasPerThread));
  const bool short_spans = nlong < 2048;
  for (int r = 0; r < nlat; r++) {
    const int per64 = (c->ld[r].nc + 63) / 64;
    const int span = short_spans ? std::min(kBlock * kMeasPerThread,
                                            std::max(kBlock, (per64 + kBlock - 1) / kBlock * kBlock))
                                 : kBlock * kMeasPerThread;
    c->ld[r].mp0 = (int)mblk.size();
    for (int col = 0; col < 2; col++)
      for (int t0 = 0; t0 < c->ld[r].nc; t0 += span)
        mblk.push_back(make_int4(r, col, t0, std::min(t0 + span, c->ld[r].nc)));
    c->ld[r].mp1 = (int)mblk.size();
  }
  c->nmblk = (int)mblk.size();

  int st;
  if ((st = dev_alloc(&c->X, c->pool)) || (st = dev_alloc(&c->Y, c->pool)) || (st = dev_alloc(&c->Z, c->pool)))
    return st;
  if (c->mask && (st = dev_alloc(&c->PT, c->pool))) return st;
  if (c->surf && ((st = dev_alloc(&c->SX, c->pool)) || (st = dev_alloc(&c->SY, c->pool)) ||
                  (st = dev_alloc(&c->SZ, c->pool))))
    return st;
  if ((st = dev_alloc(&c->surf_d, surf_h.size())) || (st = dev_alloc(&c->lat_d, nlat)) ||
      (st = dev_alloc(&c->ctrl_d, nlat)) || (st = dev_alloc(&c->sweep_base, 1)) ||
      (st = dev_alloc(&c->cnt, (size_t)nlat * kCntLattice)) ||
      (st = dev_alloc(&c->mblk_d, mblk.size())) || (st = dev_alloc(&c->partials, mblk.size() * kNQ)) ||
      (st = dev_alloc(&c->mom_d, nlat)) || (st = dev_alloc(&c->obs_d, nlat)) ||
      (st = dev_alloc(&c->scratch_f, (size_t)c->max_nt)))
    return st;
  if (!surf_h.empty())
    MCL_HIP(hipMemcpyAsync(c->surf_d, surf_h.data(), sizeof(SurfDev) * surf_h.size(), hipMemcpyHostToDevice,
                           c->stream));
  const unsigned long long zero_base = 0ull;
  MCL_HIP(hipMemcpyAsync(c->sweep_base, &zero_base, sizeof(zero_base), hipMemcpyHostToDevice, c->stream));
  // the measurement is folded into the sweep unless the pear observables are measured, some lattice measures both
  // radial orders, some lattice is a single column in x, or some lattice of a GHRL context has empty sites (each
  // colour-0 site is counted by an x neighbour; an empty one of an LL context still counts them)
  c->fused = c->pot != MCL_POT_PEAR;
  bool any_radial = false;
  for (int r = 0; r < nlat; r++) {
    if ((c->ld[r].flags & F_RSPH) && (c->ld[r].flags & F_RCYL)) c->fused = false;
    if (radial_flags(c->ld[r].flags)) any_radial = true;
  }
  c->row = true;
  c->zero_empty = c->pot == MCL_POT_LL && c->mask;
  c->resident_ok = true;
  c->resident_one = !c->flip;
  for (int r = 0; r < nlat; r++) {
    if (2ll * c->ld[r].nc > (c->flip ? kResMaxFlip : kResMax)) c->resident_ok = false;
    if (c->ld[r].nc > kResBlock) c->resident_one = false;
    if (c->ld[r].nx < 2) c->fused = false;
    if ((long long)c->ld[r].nvalid != (long long)c->ld[r].nx * c->ld[r].ny * c->ld[r].nz) {
      if (!c->zero_empty) c->fused = false;
      c->has_empty = true;
    }
    if (c->ld[r].nx % 128 != 0) c->row = false;
  }
  c->fr = c->fused && any_radial;
  if (c->fr) c->row = false;  // the radial order is finalized by the generic instances only
  // Acceptance counters in use per lattice (kc): a workgroup adds into counter (its index mod kc), a colour-0 pass
  // sums kc of them, and the sums do not depend on kc. In masked contexts that are not row-aligned (the sphere
  // batches) every colour-0 wave reads them, so a lattice uses about one per 8192 positions per colour (one for 16^3,
  // 16 for 64^3, all kCnt from 128^3 on): the 512-lattice batch +1.9 %. Every other instance takes all kCnt as a
  // constant: in unmasked contexts one wave per workgroup reads them (fewer counters were 0.4-2 % slower on single LL
  // lattices, more workgroups per counter), and the run-time count cost the masked row-aligned slab 2 %.
  for (LatDev &L : c->ld) {
    L.kc = kCnt;
    if (c->mask && !c->row) {
      L.kc = 1;
      while (L.kc < kCnt && (long long)L.kc * 8192 < (long long)L.nc) L.kc *= 2;
    }
  }
  // the lattice records, complete now
  MCL_HIP(hipMemcpyAsync(c->lat_d, c->ld.data(), sizeof(LatDev) * nlat, hipMemcpyHostToDevice, c->stream));
  {
    long long npos = 0;
    for (int r = 0; r < nlat; r++) npos += c->ld[r].nc;
    c->split = !c->flip && !c->row && npos <= (1ll << 18);
  }
  // default launch shape (before any autotuning): 512 threads in dispatch order for a single lattice, 256 threads
  // in XCD-contiguous tiles for a batch, the faster shapes measured for the 128^3 benchmark and the 512-lattice batch.
  // A single lattice too small for 1024 workgroups of 512 threads takes smaller workgroups (down to 64 threads), so
  // that its work spreads over more of the 256 CUs. Deterministic: the same lattice always gets the same shape.
  int bs0 = 256;
  if (nlat == 1) {
    const long long span = c->flip || c->split ? 1 : 2;
    bs0 = 512;
    while (bs0 > 64 && (long long)c->ld[0].nc < 1024ll * span * bs0) bs0 >>= 1;
  }
  if ((st = build_worklist(c, bs0, nlat == 1 ? 0 : 1))) return st;
  MCL_HIP(hipMemcpyAsync(c->mblk_d, mblk.data(), sizeof(int4) * mblk.size(), hipMemcpyHostToDevice, c->stream));
  std::vector<CtrlDev> ctrl(nlat);
  for (int r = 0; r < nlat; r++) {
    if (!(lat[r].T >= 0.0f)) return fail(MCL_ERR_ARG, "negative temperature");
    ctrl[r] = CtrlDev{lat[r].T, lat[r].T > 0.0f ? 1.0f / lat[r].T : 0.0f, lat[r].model.elecE, lat[r].ang_var,
                      {lat[r].ang_var, lat[r].ang_var}, 0ull};
  }
  MCL_HIP(hipMemcpyAsync(c->ctrl_d, ctrl.data(), sizeof(CtrlDev) * nlat, hipMemcpyHostToDevice, c->stream));
  std::vector<unsigned int> zeros((size_t)nlat * kCntLattice, 0u);
  MCL_HIP(hipMemcpyAsync(c->cnt, zeros.data(), sizeof(unsigned int) * zeros.size(), hipMemcpyHostToDevice,
                         c->stream));
  std::vector<MomDev> mom(nlat, MomDev{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0});
  MCL_HIP(hipMemcpyAsync(c->mom_d, mom.data(), sizeof(MomDev) * nlat, hipMemcpyHostToDevice, c->stream));
  // per-lattice point types and anchoring vectors in the colour-split layout
  for (int r = 0; r < nlat; r++) {
    const mcl_lattice_desc &d = lat[r];
    const LatHost &h = c->lh[r];
    if (c->mask) {
      std::vector<uint8_t> p(2 * (size_t)h.nc, 0);
      for (int k = 0; k < h.nz; k++)
        for (int j = 0; j < h.ny; j++)
          for (int i = 0; i < h.nx; i++) {
            const size_t lin = (size_t)i + (size_t)h.nx * (j + (size_t)h.ny * k);
            const size_t t = (size_t)((i + j + k) & 1) * h.nc + (i >> 1) + (size_t)h.nxh * (j + (size_t)h.ny * k);
            p[t] = d.pt ? (uint8_t)d.pt[lin] : (uint8_t)1;
          }
      MCL_HIP(hipMemcpy(c->PT + h.off, p.data(), p.size(), hipMemcpyHostToDevice));
    }
    if (c->surf) {
      // per-site vector of the anchoring kernels: the normal, or the easy axis for rp and rp_tangent
      // (evolve.cu:348-391)
      const long long Nt = (long long)h.nx * h.ny * h.nz;
      std::vector<float> sv(3 * (size_t)Nt, 0.0f);
      if (h.surf) {
        for (long long s = 0; s < Nt; s++) {
          const int p = d.pt[s];
          if (p < 2) continue;
          const mcl_surface &S = d.surf[p - 2];
          const float *nrm = &d.ns[3 * s];
          float *o = &sv[3 * s];
          if (S.type == MCL_ANCH_RP) {
            o[0] = S.axis[0];
            o[1] = S.axis[1];
            o[2] = S.axis[2];
          } else if (S.type == MCL_ANCH_RP_TANGENT) {
            const float ds = S.axis[0] * nrm[0] + S.axis[1] * nrm[1] + S.axis[2] * nrm[2];
            float e[3] = {S.axis[0] - ds * nrm[0], S.axis[1] - ds * nrm[1], S.axis[2] - ds * nrm[2]};
            const float mod = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (mod < 1e-3f) {  // RP_Tangent_Anchoring::singular_tol: no anchoring at this site
              o[0] = o[1] = o[2] = 0.0f;
            } else {
              o[0] = e[0] / mod;
              o[1] = e[1] / mod;
              o[2] = e[2] / mod;
            }
          } else {
            o[0] = nrm[0];
            o[1] = nrm[1];
            o[2] = nrm[2];
          }
        }
      }
      std::vector<float> x, y, z;
      aos_to_soa(h, sv.data(), x, y, z);
      const size_t bytes = sizeof(float) * 2 * (size_t)h.nc;
      MCL_HIP(hipMemcpy(c->SX + h.off, x.data(), bytes, hipMemcpyHostToDevice));
      MCL_HIP(hipMemcpy(c->SY + h.off, y.data(), bytes, hipMemcpyHostToDevice));
      MCL_HIP(hipMemcpy(c->SZ + h.off, z.data(), bytes, hipMemcpyHostToDevice));
    }
    if (c->zero_empty && d.pt && h.nvalid < (long long)h.nx * h.ny * h.nz) {
      LatHost &hw = c->lh[r];
      const long long Nt = (long long)h.nx * h.ny * h.nz;
      for (long long q = 0; q < Nt; q++)
        if (d.pt[q] == 0) hw.empty.push_back((uint32_t)q);
    }
    if ((st = upload_directors(c, r, d.ni))) return st;
  }
  MCL_HIP(hipStreamSynchronize(c->stream));
  return MCL_OK;
}


// Measurement of the lattices of group gr (c->all: every lattice); advance > 0: its finalize also advances the device
// sweep base by that much.
int run_measure(Ctx *c, int accumulate, const LatGroup &gr, unsigned advance = 0) {
  const MeasFn mf = pick_meas(*c);
  mf(dim3(gr.m1 - gr.m0), dim3(kBlock), c->stream, *c, gr.m0);
  hipLaunchKernelGGL(k_finalize, dim3(gr.l1 - gr.l0), dim3(kFinBlock), 0, c->stream, c->lat_d, c->partials, c->nmblk,
                     c->xq ? 1 : 0, c->mom_d, c->obs_d, accumulate, (const int2 *)nullptr, gr.l0, c->sweep_base,
                     advance);
  MCL_HIP(hipGetLastError());
  return MCL_OK;
}

template <int POT, bool MASK, bool SURF, bool FIELD>
void launch_site_energies(const Ctx *c, int r, const int *sites, const float *trial, int n, float *ec,
                                 float *et) {
  hipLaunchKernelGGL((k_site_energies<POT, MASK, SURF, FIELD>), dim3((n + 255) / 256), dim3(256), 0, c->stream,
                     c->lat_d, c->ctrl_d, r, sites, trial, n, c->X, c->Y, c->Z, c->PT, c->SX, c->SY, c->SZ,
                     c->surf_d, ec, et);
}
template <int POT>
void dispatch_site_energies(const Ctx *c, int r, const int *sites, const float *trial, int n, float *ec,
                                   float *et) {
  if (c->surf) {
    if (c->field) launch_site_energies<POT, true, true, true>(c, r, sites, trial, n, ec, et);
    else launch_site_energies<POT, true, true, false>(c, r, sites, trial, n, ec, et);
  } else if (c->mask) {
    if (c->field) launch_site_energies<POT, true, false, true>(c, r, sites, trial, n, ec, et);
    else launch_site_energies<POT, true, false, false>(c, r, sites, trial, n, ec, et);
  } else {
    if (c->field) launch_site_energies<POT, false, false, true>(c, r, sites, trial, n, ec, et);
    else launch_site_energies<POT, false, false, false>(c, r, sites, trial, n, ec, et);
  }
}

int set_ctrl_float(Ctx *c, int32_t lat, size_t field_off, float v) {
  if (!c) return fail(MCL_ERR_ARG, "null context");
  if (lat >= c->nlat) return fail(MCL_ERR_ARG, "lattice index out of range");
  const int r0 = lat < 0 ? 0 : lat, r1 = lat < 0 ? c->nlat : lat + 1;
  for (int r = r0; r < r1; r++) {
    char *dst = reinterpret_cast<char *>(c->ctrl_d + r) + field_off;
    MCL_HIP(hipMemcpyAsync(dst, &v, sizeof(float), hipMemcpyHostToDevice, c->stream));
  }
  MCL_HIP(hipStreamSynchronize(c->stream));
  return MCL_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// C ABI

extern "C" {

int mcl_abi_version(void) { return MCL_ABI_VERSION; }

const char *mcl_last_error(void) { return g_last_error.c_str(); }

int mcl_create(const mcl_lattice_desc *lat, int32_t nlat, void **ctx_out) {
  if (!ctx_out) return fail(MCL_ERR_ARG, "null output pointer");
  *ctx_out = nullptr;
  Ctx *c = new Ctx();
  const int st = create_ctx(lat, nlat, c);
  if (st != MCL_OK) {
    const std::string msg = g_last_error;
    destroy_ctx(c);
    g_last_error = msg;
    return st;
  }
  *ctx_out = c;
  return MCL_OK;
}

void mcl_destroy(void *ctx) { destroy_ctx(as_ctx(ctx)); }

// Queues sweeps [k0, k0 + n) relative to the device sweep base (no timing, no advance).
// The finalize of a measurement taken by the colour-1 sweep pass.
static int run_fused_finalize(Ctx *c, const LatGroup &gr, unsigned advance) {
  hipLaunchKernelGGL(k_finalize, dim3(gr.l1 - gr.l0), dim3(kFinBlock), 0, c->stream, c->lat_d, c->spart, c->nblk,
                     c->fr ? 1 : 0, c->mom_d, c->obs_d, 1, (const int2 *)c->sr_d, gr.l0, c->sweep_base, advance);
  MCL_HIP(hipGetLastError());
  return MCL_OK;
}

// One sweep (both colour passes) and, when measure is set, its measurement: folded into the colour-1 pass when the
// context allows it (its finalize then runs in the extra workgroups of the next colour-0 pass, *pending, or in a
// k_finalize launch at the end of the call), else the measurement kernels after the pass.
// *advance (when not 0) is what the device sweep base still has to gain at the end of the queued work: the
// measurement of the call's last sweep takes it into its finalize and clears it, so no k_advance launch follows.
static int sweep_once(Ctx *c, SweepFn sf, const LatGroup &gr, dim3 grid, dim3 block, unsigned k, bool measure,
                      bool *pending, bool last, unsigned *advance) {
  sf(grid, block, c->stream, *c, gr, 0, k, false, *pending ? 1 : 0);
  *pending = false;
  sf(grid, block, c->stream, *c, gr, 1, k, measure && c->fused, 0);
  if (!measure) return MCL_OK;
  if (c->fused) {
    *pending = true;
    return MCL_OK;
  }
  const unsigned adv = last ? *advance : 0u;
  if (last) *advance = 0;
  return run_measure(c, 1, gr, adv);
}

static int flush_finalize(Ctx *c, const LatGroup &gr, bool *pending, bool last, unsigned *advance) {
  if (!*pending) return MCL_OK;
  *pending = false;
  const unsigned adv = last ? *advance : 0u;
  if (last) *advance = 0;
  return run_fused_finalize(c, gr, adv);
}

// Advances the device sweep base unless the call's last finalize already did.
static int finish_advance(Ctx *c, unsigned advance) {
  if (advance) hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, c->sweep_base, advance);
  return MCL_OK;
}

// The sweeps [k0, k0 + n), lattice group after lattice group, then the advance of the sweep base by `advance`.
static int enqueue_sweeps(Ctx *c, SweepFn sf, int k0, int n, int measure_every, unsigned advance) {
  const dim3 block(c->bs);
  for (size_t g = 0; g < c->groups.size(); g++) {
    const LatGroup &gr = c->groups[g];
    const bool last_group = g + 1 == c->groups.size();
    const dim3 grid = sweep_grid(c, gr);
    bool pending = false;
    for (int s = 0; s < n; s++) {
      const int st = sweep_once(c, sf, gr, grid, block, (unsigned)(k0 + s),
                                measure_every > 0 && (s + 1) % measure_every == 0, &pending,
                                last_group && s == n - 1, &advance);
      if (st) return st;
    }
    if (int st = flush_finalize(c, gr, &pending, last_group, &advance)) return st;
  }
  return finish_advance(c, advance);
}

// Captured graphs a context keeps; the least recently replayed one is dropped beyond this.
constexpr size_t kMaxGraphs = 8;

// Graph of `len` sweeps (with the measurements of measure_every) followed by the advance of the sweep base, from the
// context's cache (most recently used last) or captured now.
static int get_graph(Ctx *c, SweepFn sf, int len, int measure_every, hipGraphExec_t *out) {
  for (size_t i = 0; i < c->graphs.size(); i++)
    if (c->graphs[i].first.first == len && c->graphs[i].first.second == measure_every) {
      std::rotate(c->graphs.begin() + i, c->graphs.begin() + i + 1, c->graphs.end());
      *out = c->graphs.back().second;
      return MCL_OK;
    }
  if (c->graphs.size() >= kMaxGraphs) {
    // the oldest graph may still be queued on the stream: its replays end before it is destroyed
    MCL_HIP(hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(c->graphs.front().second);
    c->graphs.erase(c->graphs.begin());
  }
  hipGraph_t graph = nullptr;
  MCL_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int st = enqueue_sweeps(c, sf, 0, len, measure_every, (unsigned)len);
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e2));
  c->graphs.push_back({{len, measure_every}, exec});
  *out = exec;
  return MCL_OK;
}

// How mcl_sweeps splits a call of nsweeps when graphs are on: nblocks replays of a graph of `block` sweeps (the graph
// block rounded to a multiple of measure_every), then one replay of a graph of `tail` sweeps (the largest multiple of
// measure_every in the rest, or the whole rest when nothing is measured); what remains (fewer than measure_every
// sweeps, none of them measured) runs eagerly. Every graph length is a multiple of measure_every, so the
// measurements fall where the eager path puts them.
struct GraphPlan {
  int block = 0, nblocks = 0, tail = 0;
};

static GraphPlan graph_plan(const Ctx *c, int nsweeps, int measure_every) {
  GraphPlan g;
  if (c->graph_block <= 0) return g;
  g.block = c->graph_block;
  if (measure_every > 0) g.block = measure_every * std::max(1, (c->graph_block + measure_every / 2) / measure_every);
  g.nblocks = nsweeps / g.block;
  const int rest = nsweeps - g.nblocks * g.block;
  g.tail = measure_every > 0 ? rest - rest % measure_every : rest;
  return g;
}

// Whether plain sweeps of this call run as resident launches (small lattices, no measurement, no kernel timing).
static bool runs_resident(const Ctx *c, int measure_every) {
  return measure_every == 0 && c->resident_ok && c->resident_on && c->timing_stride == 0;
}

// Replays graph g of len sweeps; with kernel timing on, one event pair around the replay.
static int replay(Ctx *c, hipGraphExec_t g, int len) {
  const bool timed = c->timing_stride > 0 && c->ev_used + 2 <= c->ev.size();
  if (timed) MCL_HIP(hipEventRecord(c->ev[c->ev_used], c->stream));
  MCL_HIP(hipGraphLaunch(g, c->stream));
  if (timed) {
    MCL_HIP(hipEventRecord(c->ev[c->ev_used + 1], c->stream));
    c->ev_runs.push_back({c->ev_used, 2.0 * len});
    c->ev_used += 2;
  }
  c->sweeps_done += (unsigned long long)len;
  return MCL_OK;
}

int mcl_prepare_sweeps(void *ctx, int32_t nsweeps, int32_t measure_every) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  if (nsweeps < 0 || measure_every < 0) return fail(MCL_ERR_ARG, "negative sweep count");
  if (nsweeps == 0 || runs_resident(c, measure_every)) return MCL_OK;
  const SweepFn sf = pick_sweep(*c);
  const GraphPlan plan = graph_plan(c, nsweeps, measure_every);
  hipGraphExec_t g;
  if (plan.nblocks > 0)
    if (int st = get_graph(c, sf, plan.block, measure_every, &g)) return st;
  if (plan.tail > 0)
    if (int st = get_graph(c, sf, plan.tail, measure_every, &g)) return st;
  return MCL_OK;
}

int mcl_sweeps(void *ctx, int32_t nsweeps, int32_t measure_every) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  if (nsweeps < 0 || measure_every < 0) return fail(MCL_ERR_ARG, "negative sweep count");
  if (nsweeps == 0) return MCL_OK;
  if (runs_resident(c, measure_every)) {
    // plain sweeps of small lattices: whole sweeps in one launch per kResChunk, bit-identical to the passes below
    const ResFn rf = pick_resident(*c);
    for (int k0 = 0; k0 < nsweeps; k0 += kResChunk) rf(c->stream, *c, (unsigned)k0, std::min(kResChunk, nsweeps - k0));
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, c->stream, c->sweep_base, (unsigned)nsweeps);
    c->sweeps_done += (unsigned long long)nsweeps;
    MCL_HIP(hipGetLastError());
    return MCL_OK;
  }
  const SweepFn sf = pick_sweep(*c);
  const dim3 block(c->bs);
  int done = 0;
  const GraphPlan plan = graph_plan(c, nsweeps, measure_every);
  if (plan.nblocks > 0) {
    hipGraphExec_t g;
    if (int st = get_graph(c, sf, plan.block, measure_every, &g)) return st;
    for (int b = 0; b < plan.nblocks; b++, done += plan.block)
      if (int st = replay(c, g, plan.block)) return st;
  }
  if (plan.tail > 0) {
    hipGraphExec_t g;
    if (int st = get_graph(c, sf, plan.tail, measure_every, &g)) return st;
    if (int st = replay(c, g, plan.tail)) return st;
    done += plan.tail;
  }
  const int rest = nsweeps - done;
  if (rest > 0) {
    // Kernel timing: one event pair around each run of consecutive sweep launches of a lattice group (up to
    // timing_stride sweeps, cut where separate measurement kernels follow a sweep), so that the markers do not sit
    // between the timed kernels and the measurement kernels are not counted as sweep time. A measurement folded into
    // the sweep adds no kernel, so it does not cut the run.
    unsigned advance = (unsigned)rest;
    for (size_t g = 0; g < c->groups.size(); g++) {
      const LatGroup &gr = c->groups[g];
      const bool last_group = g + 1 == c->groups.size();
      const dim3 grid = sweep_grid(c, gr);
      int run_len = 0;
      size_t run_start = 0;
      bool pending = false;
      for (int s = 0; s < rest; s++) {
        if (c->timing_stride > 0 && run_len == 0 && c->ev_used + 2 <= c->ev.size()) {
          run_start = c->ev_used;
          c->ev_used += 2;
          MCL_HIP(hipEventRecord(c->ev[run_start], c->stream));
          run_len = 1;
        } else if (run_len > 0) {
          run_len++;
        }
        const bool measure_now = measure_every > 0 && (s + 1) % measure_every == 0;
        sf(grid, block, c->stream, *c, gr, 0, (unsigned)s, false, pending ? 1 : 0);
        pending = false;
        sf(grid, block, c->stream, *c, gr, 1, (unsigned)s, measure_now && c->fused, 0);
        if (run_len > 0 && ((measure_now && !c->fused) || run_len == c->timing_stride || s == rest - 1)) {
          MCL_HIP(hipEventRecord(c->ev[run_start + 1], c->stream));
          c->ev_runs.push_back({run_start, 2.0 * run_len * gr.share});
          run_len = 0;
        }
        if (measure_now) {
          if (c->fused) {
            pending = true;
          } else {
            const bool last = last_group && s == rest - 1;
            const int st = run_measure(c, 1, gr, last ? advance : 0u);
            if (last) advance = 0;
            if (st) return st;
          }
        }
      }
      if (int st = flush_finalize(c, gr, &pending, last_group, &advance)) return st;
    }
    if (int st = finish_advance(c, advance)) return st;
    c->sweeps_done += (unsigned long long)rest;
  }
  MCL_HIP(hipGetLastError());
  return MCL_OK;
}

int mcl_set_kernel_timing(void *ctx, int32_t stride, int32_t max_timed_sweeps) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  MCL_HIP(hipStreamSynchronize(c->stream));
  c->timing_stride = stride > 0 ? stride : 0;
  c->ev_used = 0;
  c->ev_runs.clear();
  // create the events up front so that no hipEventCreate falls inside a timed region
  while (c->ev.size() < 2 * (size_t)std::max(0, max_timed_sweeps)) {
    hipEvent_t e;
    MCL_HIP(hipEventCreate(&e));
    c->ev.push_back(e);
  }
  return MCL_OK;
}

int mcl_kernel_timing(void *ctx, int64_t *launches, double *total_ms) {
  Ctx *c = as_ctx(ctx);
  if (!c || !launches || !total_ms) return fail(MCL_ERR_ARG, "null argument");
  MCL_HIP(hipStreamSynchronize(c->stream));
  double tot = 0.0, n = 0.0;
  for (auto &p : c->ev_runs) {
    float ms = 0.0f;
    MCL_HIP(hipEventElapsedTime(&ms, c->ev[p.first], c->ev[p.first + 1]));
    tot += ms;
    n += p.second;
  }
  *launches = (int64_t)std::llround(n);
  *total_ms = tot;
  return MCL_OK;
}

int mcl_get_stream(void *ctx, void **stream) {
  Ctx *c = as_ctx(ctx);
  if (!c || !stream) return fail(MCL_ERR_ARG, "null argument");
  *stream = (void *)c->stream;
  return MCL_OK;
}

int mcl_set_temperature(void *ctx, int32_t lat, float T) {
  if (!(T >= 0.0f)) return fail(MCL_ERR_ARG, "negative temperature");
  Ctx *c = as_ctx(ctx);
  int st = set_ctrl_float(c, lat, offsetof(CtrlDev, T), T);
  if (st) return st;
  return set_ctrl_float(c, lat, offsetof(CtrlDev, invT), T > 0.0f ? 1.0f / T : 0.0f);
}
int mcl_set_field(void *ctx, int32_t lat, float elecE) {
  return set_ctrl_float(as_ctx(ctx), lat, offsetof(CtrlDev, elecE), elecE);
}

int mcl_set_ang_var(void *ctx, int32_t lat, float ang_var) {
  Ctx *c = as_ctx(ctx);
  int st = set_ctrl_float(c, lat, offsetof(CtrlDev, av_set), ang_var);
  if (st) return st;
  const int r0 = lat < 0 ? 0 : lat, r1 = lat < 0 ? c->nlat : lat + 1;
  const unsigned long long sw = c->sweeps_done;
  for (int r = r0; r < r1; r++)
    MCL_HIP(hipMemcpyAsync(reinterpret_cast<char *>(c->ctrl_d + r) + offsetof(CtrlDev, av_set_sweep), &sw,
                           sizeof(sw), hipMemcpyHostToDevice, c->stream));
  MCL_HIP(hipStreamSynchronize(c->stream));
  return MCL_OK;
}

// Accepted rotations of the last completed sweep and the ang_var that the next sweep will use.
static int sweep_state(Ctx *c, int lat, unsigned *last_acc, float *next_av) {
  CtrlDev h;
  std::vector<unsigned> v(kCntSweep, 0u);
  MCL_HIP(hipMemcpyAsync(&h, c->ctrl_d + lat, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  const unsigned long long S = c->sweeps_done;
  unsigned a = 0;
  if (S > 0) {
    MCL_HIP(hipMemcpyAsync(v.data(), c->cnt + (size_t)lat * kCntLattice + ((S - 1) % 3) * kCntSweep,
                           sizeof(unsigned) * kCntSweep, hipMemcpyDeviceToHost, c->stream));
  }
  MCL_HIP(hipStreamSynchronize(c->stream));
  for (int x = 0; x < kCnt; x++) a += v[(size_t)x * kCntStride];
  *last_acc = a;
  if (S == h.av_set_sweep)
    *next_av = h.av_set;
  else
    *next_av = ang_var_rule(h.av_store[(S - 1) & 1], a, c->lh[lat].nvalid);
  return MCL_OK;
}

int mcl_get_ang_var(void *ctx, int32_t lat, float *ang_var) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  if (!ang_var) return fail(MCL_ERR_ARG, "null argument");
  unsigned a;
  return sweep_state(c, lat, &a, ang_var);
}

int mcl_get_sweep_stats(void *ctx, int32_t lat, int64_t *last_accepted, int64_t *sweeps_done) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  unsigned a;
  float av;
  if ((st = sweep_state(c, lat, &a, &av))) return st;
  if (last_accepted) *last_accepted = (int64_t)a;
  if (sweeps_done) *sweeps_done = (int64_t)c->sweeps_done;
  return MCL_OK;
}

int mcl_reset_moments(void *ctx) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  hipLaunchKernelGGL(k_zero_moments, dim3((c->nlat + 255) / 256), dim3(256), 0, c->stream, c->mom_d, c->nlat);
  MCL_HIP(hipGetLastError());
  return MCL_OK;
}

int mcl_get_moments(void *ctx, int32_t lat, mcl_moments *out) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  if (!out) return fail(MCL_ERR_ARG, "null argument");
  MomDev m;
  MCL_HIP(hipMemcpyAsync(&m, c->mom_d + lat, sizeof(m), hipMemcpyDeviceToHost, c->stream));
  MCL_HIP(hipStreamSynchronize(c->stream));
  out->n = (int64_t)m.n;
  out->S = m.S;
  out->S2 = m.S2;
  out->E = m.E;
  out->E2 = m.E2;
  out->Sr_sph = m.srs;
  out->Pr_sph = m.prs;
  out->Sr_cyl = m.src;
  out->Pr_cyl = m.prc;
  out->P = m.P;
  out->dtheta = m.dth;
  out->divn = m.divn;
  return MCL_OK;
}

int mcl_measure(void *ctx, int32_t lat, mcl_observables *out) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  if (!out) return fail(MCL_ERR_ARG, "null argument");
  if ((st = run_measure(c, 0, c->all))) return st;
  ObsDev o;
  MCL_HIP(hipMemcpyAsync(&o, c->obs_d + lat, sizeof(o), hipMemcpyDeviceToHost, c->stream));
  MCL_HIP(hipStreamSynchronize(c->stream));
  out->nvalid = (int64_t)o.nvalid;
  out->E = o.E;
  for (int a = 0; a < 6; a++) out->Q[a] = o.Q[a];
  out->S = o.S;
  for (int a = 0; a < 3; a++) out->director[a] = o.dir[a];
  out->Sr_sph = o.srs;
  out->Pr_sph = o.prs;
  out->Sr_cyl = o.src;
  out->Pr_cyl = o.prc;
  out->P = o.P;
  out->dtheta = o.dth;
  out->divn = o.divn;
  return MCL_OK;
}

int mcl_get_directors(void *ctx, int32_t lat, float *ni) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  if (!ni) return fail(MCL_ERR_ARG, "null argument");
  const LatHost &h = c->lh[lat];
  const size_t n = 2 * (size_t)h.nc;
  std::vector<float> x(n), y(n), z(n);
  MCL_HIP(hipMemcpyAsync(x.data(), c->X + h.off, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  MCL_HIP(hipMemcpyAsync(y.data(), c->Y + h.off, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  MCL_HIP(hipMemcpyAsync(z.data(), c->Z + h.off, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  MCL_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < h.nz; k++)
    for (int j = 0; j < h.ny; j++)
      for (int i = 0; i < h.nx; i++) {
        const size_t lin = (size_t)i + (size_t)h.nx * (j + (size_t)h.ny * k);
        const size_t t = (size_t)((i + j + k) & 1) * h.nc + (i >> 1) + (size_t)h.nxh * (j + (size_t)h.ny * k);
        ni[3 * lin] = x[t];
        ni[3 * lin + 1] = y[t];
        ni[3 * lin + 2] = z[t];
      }
  for (size_t e = 0; e < h.empty.size() && c->zero_empty; e++)
    for (int a = 0; a < 3; a++) ni[3 * (size_t)h.empty[e] + a] = h.empty_ni[3 * e + a];
  return MCL_OK;
}

int mcl_set_directors(void *ctx, int32_t lat, const float *ni) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  if (!ni) return fail(MCL_ERR_ARG, "null argument");
  return upload_directors(c, lat, ni);
}

int mcl_local_order(void *ctx, int32_t lat, float *out) {
  Ctx *c = as_ctx(ctx);
  int st = check_lat(c, lat);
  if (st) return st;
  if (!out) return fail(MCL_ERR_ARG, "null argument");
  const LatHost &h = c->lh[lat];
  const int n = 2 * h.nc;
  const dim3 grid((n + kBlock - 1) / kBlock);
  if (c->mask)
    hipLaunchKernelGGL(k_local_order<true>, grid, dim3(kBlock), 0, c->stream, c->lat_d, lat, c->X, c->Y, c->Z,
                       c->PT, c->scratch_f);
  else
    hipLaunchKernelGGL(k_local_order<false>, grid, dim3(kBlock), 0, c->stream, c->lat_d, lat, c->X, c->Y, c->Z,
                       c->PT, c->scratch_f);
  MCL_HIP(hipGetLastError());
  const size_t Nt = (size_t)h.nx * h.ny * h.nz;
  MCL_HIP(hipMemcpyAsync(out, c->scratch_f, sizeof(float) * Nt, hipMemcpyDeviceToHost, c->stream));
  MCL_HIP(hipStreamSynchronize(c->stream));
  return MCL_OK;
}

int mcl_synchronize(void *ctx) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  MCL_HIP(hipStreamSynchronize(c->stream));
  return MCL_OK;
}

// Checks of the site-energy calls: the lattice, the pointers and every site index.
static int check_site_query(Ctx *c, int32_t lat, const int32_t *sites, const float *trial, int32_t n,
                            const float *e_cur, const float *e_trial) {
  int st = check_lat(c, lat);
  if (st) return st;
  if (n <= 0) return MCL_OK;
  if (!sites || !trial || !e_cur || !e_trial) return fail(MCL_ERR_ARG, "null argument");
  const LatHost &h = c->lh[lat];
  const long long Nt = (long long)h.nx * h.ny * h.nz;
  for (int g = 0; g < n; g++)
    if (sites[g] < 0 || sites[g] >= Nt) return fail(MCL_ERR_ARG, "site index out of range");
  return MCL_OK;
}

// Site energies through device staging buffers: copies in, the sweep's energy function, copies out, wait.
static hipError_t site_energies_staged(Ctx *c, int32_t lat, const int32_t *sites, const float *trial, int32_t n,
                                       float *e_cur, float *e_trial, int *d_sites, float *d_trial, float *d_ec,
                                       float *d_et) {
  hipError_t e = hipMemcpyAsync(d_sites, sites, sizeof(int) * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_trial, trial, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    switch (c->pot) {
      case MCL_POT_GHRL: dispatch_site_energies<MCL_POT_GHRL>(c, lat, d_sites, d_trial, n, d_ec, d_et); break;
      case MCL_POT_PEAR: dispatch_site_energies<MCL_POT_PEAR>(c, lat, d_sites, d_trial, n, d_ec, d_et); break;
      default: dispatch_site_energies<MCL_POT_LL>(c, lat, d_sites, d_trial, n, d_ec, d_et); break;
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(e_cur, d_ec, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(e_trial, d_et, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

int mcl_site_energies(void *ctx, int32_t lat, const int32_t *sites, const float *trial, int32_t n, float *e_cur,
                      float *e_trial) {
  Ctx *c = as_ctx(ctx);
  int st = check_site_query(c, lat, sites, trial, n, e_cur, e_trial);
  if (st || n <= 0) return st;
  int *d_sites = nullptr;
  float *d_trial = nullptr, *d_ec = nullptr, *d_et = nullptr;
  if ((st = dev_alloc(&d_sites, n)) || (st = dev_alloc(&d_trial, 3 * (size_t)n)) || (st = dev_alloc(&d_ec, n)) ||
      (st = dev_alloc(&d_et, n))) {
    (void)hipFree(d_sites);
    (void)hipFree(d_trial);
    (void)hipFree(d_ec);
    (void)hipFree(d_et);
    return st;
  }
  const hipError_t e = site_energies_staged(c, lat, sites, trial, n, e_cur, e_trial, d_sites, d_trial, d_ec, d_et);
  (void)hipFree(d_sites);
  (void)hipFree(d_trial);
  (void)hipFree(d_ec);
  (void)hipFree(d_et);
  if (e != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("site energies: ") + hipGetErrorString(e));
  return MCL_OK;
}

int mcl_site_energies_noalloc(void *ctx, int32_t lat, const int32_t *sites, const float *trial, int32_t n,
                              float *e_cur, float *e_trial) {
  Ctx *c = as_ctx(ctx);
  if (c && n > MCL_SCRATCH_SITES)
    return fail(MCL_ERR_ARG, "site energies without allocation: " + std::to_string(n) + " sites requested, at most " +
                                 std::to_string(MCL_SCRATCH_SITES) + " per call");
  int st = check_site_query(c, lat, sites, trial, n, e_cur, e_trial);
  if (st || n <= 0) return st;
  unsigned char *base = static_cast<unsigned char *>(scratch_address(c->device));
  if (!base) return fail(MCL_ERR_DEVICE, "site energies without allocation: no scratch address for this device");
  // [site indices | trial directors | current energies | trial energies], each sized for MCL_SCRATCH_SITES
  int *d_sites = reinterpret_cast<int *>(base);
  float *d_trial = reinterpret_cast<float *>(base + sizeof(int) * (size_t)MCL_SCRATCH_SITES);
  float *d_ec = d_trial + 3 * (size_t)MCL_SCRATCH_SITES;
  float *d_et = d_ec + MCL_SCRATCH_SITES;
  const hipError_t e = site_energies_staged(c, lat, sites, trial, n, e_cur, e_trial, d_sites, d_trial, d_ec, d_et);
  if (e != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("site energies: ") + hipGetErrorString(e));
  return MCL_OK;
}

int mcl_philox(const uint32_t *ctr, const uint32_t *key, int32_t n, uint32_t *out) {
  if (n <= 0) return MCL_OK;
  if (!ctr || !key || !out) return fail(MCL_ERR_ARG, "null argument");
  uint32_t *d_ctr = nullptr, *d_key = nullptr, *d_out = nullptr;
  int st;
  if ((st = dev_alloc(&d_ctr, 4 * (size_t)n)) || (st = dev_alloc(&d_key, 2 * (size_t)n)) ||
      (st = dev_alloc(&d_out, 4 * (size_t)n))) {
    (void)hipFree(d_ctr);
    (void)hipFree(d_key);
    (void)hipFree(d_out);
    return st;
  }
  hipError_t e = hipMemcpy(d_ctr, ctr, sizeof(uint32_t) * 4 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_key, key, sizeof(uint32_t) * 2 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_philox, dim3((n + 255) / 256), dim3(256), 0, 0, d_ctr, d_key, n, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d_out, sizeof(uint32_t) * 4 * n, hipMemcpyDeviceToHost);
  (void)hipFree(d_ctr);
  (void)hipFree(d_key);
  (void)hipFree(d_out);
  if (e != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("philox: ") + hipGetErrorString(e));
  return MCL_OK;
}

int mcl_copy_bandwidth(int64_t nbytes, int32_t reps, double *gbs, double *ms_per_copy) {
  if (nbytes < 16 || reps < 1 || !gbs) return fail(MCL_ERR_ARG, "copy: bad arguments");
  const size_t n = (size_t)nbytes / 16;
  float4 *src = nullptr, *dst = nullptr;
  int st;
  if ((st = dev_alloc(&src, n)) || (st = dev_alloc(&dst, n))) {
    (void)hipFree(src);
    (void)hipFree(dst);
    return st;
  }
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipMemset(src, 0, n * sizeof(float4));
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  const dim3 grid(256 * 16);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_copy, grid, dim3(kBlock), 0, 0, src, dst, n);  // warm-up
    e = hipEventRecord(a, 0);
  }
  for (int r = 0; e == hipSuccess && r < reps; r++) {
    hipLaunchKernelGGL(k_copy, grid, dim3(kBlock), 0, 0, src, dst, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(b, 0);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (e != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("copy: ") + hipGetErrorString(e));
  const double per = (double)ms / reps;
  *gbs = 2.0 * (double)(n * sizeof(float4)) / (per * 1e6);  // bytes read + written
  if (ms_per_copy) *ms_per_copy = per;
  return MCL_OK;
}

int mcl_set_graph(void *ctx, int32_t block_sweeps) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  if (block_sweeps < 0) return fail(MCL_ERR_ARG, "negative graph block");
  c->graph_block = block_sweeps;
  return MCL_OK;
}

int mcl_set_lattice_groups(void *ctx, int64_t max_bytes) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  c->group_bytes = max_bytes > 0 ? (long long)max_bytes : 0;
  return build_worklist(c, c->bs, c->xcd);
}

int mcl_get_lattice_groups(void *ctx, int32_t *ngroups) {
  Ctx *c = as_ctx(ctx);
  if (!c || !ngroups) return fail(MCL_ERR_ARG, "null argument");
  *ngroups = (int32_t)c->groups.size();
  return MCL_OK;
}

int mcl_set_resident(void *ctx, int32_t enable) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  c->resident_on = enable != 0;
  return MCL_OK;
}

int mcl_resident(void *ctx, int32_t *active) {
  Ctx *c = as_ctx(ctx);
  if (!c || !active) return fail(MCL_ERR_ARG, "null argument");
  *active = c->resident_ok && c->resident_on ? 1 : 0;
  return MCL_OK;
}

int mcl_measure_in_sweep(void *ctx, int32_t *active) {
  Ctx *c = as_ctx(ctx);
  if (!c || !active) return fail(MCL_ERR_ARG, "null argument");
  *active = c->fused ? 1 : 0;
  return MCL_OK;
}

int mcl_set_launch(void *ctx, int32_t threads, int32_t xcd_order) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(MCL_ERR_ARG, "null context");
  if (threads < 64 || threads > 1024 || threads % 64) return fail(MCL_ERR_ARG, "threads must be 64..1024, x64");
  return build_worklist(c, threads, xcd_order ? 1 : 0);
}

}  // extern "C"

namespace {

std::string tune_key(const Ctx *c) {
  int maxnc = 0;
  for (const auto &h : c->lh) maxnc = std::max(maxnc, h.nc);
  int lp = 0, lm = 0;
  while ((1ll << lp) < c->pool) lp++;
  while ((1 << lm) < maxnc) lm++;
  char buf[192];
  // every flag that selects the kernel instance (so contexts that run different kernels do not share a result),
  // then the sizes
  snprintf(buf, sizeof(buf),
           "sweep:pot%d:mask%d:surf%d:field%d:flip%d:free%d:row%d:split%d:fused%d:empty%d:xq%d:nlat%d:pool2^%d:maxnc2^%d"
           ":groupmb%lld",
           c->pot, c->mask, c->surf, c->field, c->flip, c->free, c->row, c->split, c->fused, c->has_empty, c->xq,
           c->nlat, lp, lm, c->group_bytes > 0 ? c->group_bytes >> 20 : 0ll);
  return buf;
}

std::vector<std::pair<std::string, mcl_tune_result>> g_tune_cache;

bool tune_lookup(const std::string &key, mcl_tune_result *out) {
  for (auto &e : g_tune_cache)
    if (e.first == key) {
      *out = e.second;
      return true;
    }
  const char *path = getenv("MCL_TUNE_CACHE");
  if (!path) return false;
  FILE *f = fopen(path, "r");
  if (!f) return false;
  char k[200];
  int th, xo;
  double us;
  bool found = false;
  while (fscanf(f, "%199s %d %d %lf", k, &th, &xo, &us) == 4)
    if (key == k) {
      out->threads = th;
      out->xcd_order = xo;
      out->us_per_sweep = us;
      found = true;
    }
  fclose(f);
  if (found) g_tune_cache.push_back({key, *out});
  return found;
}

void tune_store(const std::string &key, const mcl_tune_result &r) {
  g_tune_cache.push_back({key, r});
  const char *path = getenv("MCL_TUNE_CACHE");
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %.4f\n", key.c_str(), r.threads, r.xcd_order, r.us_per_sweep);
  fclose(f);
}

}  // namespace

extern "C" {

// Autotuner of the sweep launch shape: times every (threads per workgroup, tile order) candidate on the context's
// own lattices and keeps the fastest. The lattices, control blocks, counters and sweep count are saved before the
// trials and restored after them, so tuning leaves the simulation state untouched.
int mcl_autotune(void *ctx, int32_t trial_sweeps, mcl_tune_result *out) {
  return mcl_autotune_ex(ctx, trial_sweeps, 0, out);
}

int mcl_autotune_ex(void *ctx, int32_t trial_sweeps, int32_t measure_every, mcl_tune_result *out) {
  Ctx *c = as_ctx(ctx);
  if (!c || !out) return fail(MCL_ERR_ARG, "null argument");
  if (trial_sweeps < 1) return fail(MCL_ERR_ARG, "trial_sweeps must be >= 1");
  if (measure_every < 0) return fail(MCL_ERR_ARG, "measure_every must be >= 0");
  const std::string key = tune_key(c) + ":every" + std::to_string(measure_every);
  mcl_tune_result best{};
  if (tune_lookup(key, &best)) {
    best.from_cache = 1;
    int st = build_worklist(c, best.threads, best.xcd_order);
    if (st) return st;
    *out = best;
    return MCL_OK;
  }
  MCL_HIP(hipStreamSynchronize(c->stream));
  // save the state the trials change
  float *bx = nullptr, *by = nullptr, *bz = nullptr;
  CtrlDev *bctrl = nullptr;
  unsigned *bcnt = nullptr;
  MomDev *bmom = nullptr;
  ObsDev *bobs = nullptr;
  int st;
  const size_t pool = (size_t)c->pool, ncnt = (size_t)c->nlat * kCntLattice;
  auto release = [&]() {
    (void)hipFree(bx);
    (void)hipFree(by);
    (void)hipFree(bz);
    (void)hipFree(bctrl);
    (void)hipFree(bcnt);
    (void)hipFree(bmom);
    (void)hipFree(bobs);
  };
  if ((st = dev_alloc(&bx, pool)) || (st = dev_alloc(&by, pool)) || (st = dev_alloc(&bz, pool)) ||
      (st = dev_alloc(&bctrl, c->nlat)) || (st = dev_alloc(&bcnt, ncnt)) || (st = dev_alloc(&bmom, c->nlat)) ||
      (st = dev_alloc(&bobs, c->nlat))) {
    release();
    return st;
  }
  hipError_t e = hipMemcpyAsync(bx, c->X, sizeof(float) * pool, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(by, c->Y, sizeof(float) * pool, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bz, c->Z, sizeof(float) * pool, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(bctrl, c->ctrl_d, sizeof(CtrlDev) * c->nlat, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bcnt, c->cnt, sizeof(unsigned) * ncnt, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bmom, c->mom_d, sizeof(MomDev) * c->nlat, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bobs, c->obs_d, sizeof(ObsDev) * c->nlat, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    release();
    return fail(MCL_ERR_DEVICE, std::string("autotune save: ") + hipGetErrorString(e));
  }
  const unsigned long long saved_sweeps = c->sweeps_done;
  const int saved_graph = c->graph_block, saved_timing = c->timing_stride;
  c->graph_block = 0;
  c->timing_stride = 0;
  hipEvent_t a = nullptr, b = nullptr;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const int threads[] = {64, 128, 256, 512, 1024};
  best.us_per_sweep = 1e30;
  for (int th : threads) {
    for (int xo = 0; xo < 2 && st == MCL_OK && e == hipSuccess; xo++) {
      if ((st = build_worklist(c, th, xo))) break;
      if ((st = mcl_sweeps(c, 2, 0))) break;  // warm-up
      if ((e = hipEventRecord(a, c->stream)) != hipSuccess) break;
      if ((st = mcl_sweeps(c, trial_sweeps, measure_every))) break;
      if ((e = hipEventRecord(b, c->stream)) != hipSuccess) break;
      if ((e = hipEventSynchronize(b)) != hipSuccess) break;
      float ms = 0.0f;
      if ((e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) break;
      const double us = 1e3 * ms / trial_sweeps;
      if (us < best.us_per_sweep) {
        best.us_per_sweep = us;
        best.threads = th;
        best.xcd_order = xo;
      }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  // restore the state
  const unsigned long long base = saved_sweeps;
  hipError_t r = hipMemcpyAsync(c->X, bx, sizeof(float) * pool, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess) r = hipMemcpyAsync(c->Y, by, sizeof(float) * pool, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess) r = hipMemcpyAsync(c->Z, bz, sizeof(float) * pool, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess)
    r = hipMemcpyAsync(c->ctrl_d, bctrl, sizeof(CtrlDev) * c->nlat, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess) r = hipMemcpyAsync(c->cnt, bcnt, sizeof(unsigned) * ncnt, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess) r = hipMemcpyAsync(c->mom_d, bmom, sizeof(MomDev) * c->nlat, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess) r = hipMemcpyAsync(c->obs_d, bobs, sizeof(ObsDev) * c->nlat, hipMemcpyDeviceToDevice, c->stream);
  if (r == hipSuccess) r = hipMemcpyAsync(c->sweep_base, &base, sizeof(base), hipMemcpyHostToDevice, c->stream);
  if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
  release();
  c->sweeps_done = saved_sweeps;
  c->graph_block = saved_graph;
  c->timing_stride = saved_timing;
  if (st) return st;
  if (e != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
  if (r != hipSuccess) return fail(MCL_ERR_DEVICE, std::string("autotune restore: ") + hipGetErrorString(r));
  best.from_cache = 0;
  if ((st = build_worklist(c, best.threads, best.xcd_order))) return st;
  tune_store(key, best);
  *out = best;
  return MCL_OK;
}

}  // extern "C"
