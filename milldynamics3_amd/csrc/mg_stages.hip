// One slurry-plus-media substep as a chain of stage kernels, one workgroup per drum.
//
// The chain restates FluidParticles::step_coupled (crates/mill-core/src/pbf.rs:1174-2365), then
// DemState::step_recorded (crates/mill-core/src/dem.rs:530-1066), then the Simulation
// accumulators (crates/mill-core/src/lib.rs:470-565, stress.rs:110-171):
//
//   k_fluid_prep   sanitize, predict, ball grid, fluid grid, neighbour CSR      pbf.rs:1200-1239
//   k_density      adaptive density-constraint passes                           pbf.rs:1257-1334
//   k_couple       overlap push, cohesion, adhesion                             pbf.rs:1361-1631
//   k_wall         wall projection, velocity rebuild, no-slip                   pbf.rs:1635-1700
//   k_drag         form drag (Gauss-Seidel over balls)                          pbf.rs:1777-2013
//   k_buoy         buoyancy, density recompute, viscosity weights               pbf.rs:2035-2154
//   k_links        ball boundary groups, squeeze links                          pbf.rs:2159-2241
//   k_visc         implicit viscosity CG (fluid only or joint with the balls)   pbf.rs:2244-2299
//   k_fluid_tail   shear rate, speed clamp, impulse clamp                       pbf.rs:2303-2345
//   k_dem_a        DEM predict, broadphase pairs, CCD                            dem.rs:558-706
//   k_dem_b        projection, friction, restitution, rolling, stats             dem.rs:708-1065
//   k_epilogue     energy budget, EMAs, stress accumulator, clock                lib.rs:486-565
//
// Mapping to CDNA4: blockIdx.x = drum slot, 1024 threads (16 wave64s) per drum; a batch of drums
// runs every stage in one launch. Particle-parallel stages run one particle per thread. Stages the
// reference runs Gauss-Seidel over balls keep its order: one wave walks the balls in index order
// while its lanes evaluate the candidate neighbours of the current ball, and order-sensitive sums
// are accumulated in the reference's order from the lanes' values (readlane). Sums the reference
// accumulates over a foreign index are gathered in the reference's order through an ordered CSR
// (counting sort + per-key insertion sort), never with floating-point atomics, so results are
// run-to-run deterministic. All helpers are inlined: the kernels make no device-function calls
// and pass no pointers to private memory.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_dev.h"
#include "mg_layout.h"

#define GRAV (-9.81f)
#define KERNEL extern "C" __global__ void __launch_bounds__(MG_BLOCK)

// ---------------------------------------------------------------------------------------------
// Drum SDF (geometry.rs:37-108, 233-262); the angle's sine/cosine come from the schedule.
// ---------------------------------------------------------------------------------------------
DEV float poly_sdf4(float2 p, float2 v0, float2 v1, float2 v2_, float2 v3) {
  float min_d2 = 3.4028235e38f;
  bool inside = true;
  for (int i = 0; i < 4; i++) {
    float2 a = i == 0 ? v0 : (i == 1 ? v1 : (i == 2 ? v2_ : v3));
    float2 b = i == 0 ? v1 : (i == 1 ? v2_ : (i == 2 ? v3 : v0));
    float2 edge = vsub(b, a), to_p = vsub(p, a);
    float ls = vlen2(edge);
    if (ls < 1e-12f) continue;
    float t = fclamp(vdot(to_p, edge) / ls, 0.0f, 1.0f);
    float2 closest = vadd(a, vmul(edge, t));
    min_d2 = fminf(min_d2, vlen2(vsub(p, closest)));
    float cr = edge.x * to_p.y - edge.y * to_p.x;
    if (cr < 0.0f) inside = false;
  }
  float d = sqrtf(min_d2);
  return inside ? d : -d;
}
DEV float drum_sdf(const MgDrum& D, float2 p) {
  float dist = D.R - vlen(p);
  if (D.lcount == 0) return dist;
  float max_inside = -INFINITY;
  float rt = D.R - D.l_height;
  float2 a = v2(D.R, -D.l_base_half), b = v2(D.R, D.l_base_half), c = v2(rt, D.l_top_half), d = v2(rt, -D.l_top_half);
  for (uint32_t k = 0; k < D.lcount && k < MG_MAX_LIFTERS; k++) {
    float s = D.lsin[k], co = D.lcos[k];
    float pr = p.x * co + p.y * s;
    float pt = -p.x * s + p.y * co;
    max_inside = fmaxf(max_inside, poly_sdf4(v2(pr, pt), a, b, c, d));
  }
  return fminf(dist, -max_inside);
}
// sdf_world with from_angle(-angle) = (cl, sl) and from_angle(angle) = (cw, sw).
DEV float drum_sdf_world(const MgDrum& D, float2 pw, float cl, float sl, float cw, float sw, float2& nrm) {
  if (!vfinite(pw)) { nrm = v2(0.0f, 0.0f); return 0.0f; }
  float2 pl = vrot(pw, cl, sl);
  const float EPS = 1e-4f;
  float d = drum_sdf(D, pl);
  float dx = drum_sdf(D, vadd(pl, v2(EPS, 0.0f))) - drum_sdf(D, vsub(pl, v2(EPS, 0.0f)));
  float dy = drum_sdf(D, vadd(pl, v2(0.0f, EPS))) - drum_sdf(D, vsub(pl, v2(0.0f, EPS)));
  float2 g = vdiv(v2(dx, dy), 2.0f * EPS);
  float2 nl = vlen2(g) > 1e-12f ? vnormalize(g) : vneg(vnormalize_or_zero(pl));
  nrm = vrot(nl, cw, sw);
  return d;
}
DEV float2 wall_velocity(float omega, float2 p) { return smul(omega, v2(-p.y, p.x)); }
// toi_swept (geometry.rs:146-194); from_angle(-a0) is supplied, later angles use device sin/cos.
DEV bool toi_swept(const MgDrum& D, float2 x0, float2 x1, float a0, float a1, float ca0, float sa0, float radius,
                   float& out) {
  float2 l0 = vrot(x0, ca0, sa0);
  if (drum_sdf(D, l0) - radius <= 0.0f) return false;
  float2 path = vsub(x1, x0);
  float pl = vlen(path);
  float da = a1 - a0;
  float pos_bound = vlen(x0) + pl;
  float lsb = pl + fabsf(da) * pos_bound;
  if (lsb < 1e-9f) return false;
  float t = 0.0f;
  for (int it = 0; it < 32; it++) {
    float at = a0 + da * t;
    float2 xt = vadd(x0, vmul(path, t));
    float c = it == 0 ? ca0 : cosf(-at);
    float s = it == 0 ? sa0 : sinf(-at);
    float2 loc = vrot(xt, c, s);
    float d = drum_sdf(D, loc) - radius;
    if (d <= 0.0f) { out = t; return true; }
    float st = fmaxf(d / lsb, 1e-6f);
    t += st;
    if (t >= 1.0f) return false;
  }
  out = fminf(t, 1.0f);
  return true;
}

// ---------------------------------------------------------------------------------------------
// Wave helpers for the reference-ordered per-ball loops.
// ---------------------------------------------------------------------------------------------
DEV int lane_id() { return (int)(threadIdx.x & 63); }
DEV float rlf(float v, int src) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src)); }
DEV int next_lane(uint64_t& m) {
  int s = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)m) - 1);
  m &= m - 1ull;
  return s;
}
// Number of candidates of the 3x3 query around cell (cx, cy), and candidate c by index.
DEV uint32_t near_total(const GridView& g, int32_t cx, int32_t cy) {
  uint32_t t = 0, s, e;
  for (int q = 0; q < 9; q++)
    if (near_cell(g, cx, cy, q, s, e)) t += e - s;
  return t;
}
DEV uint32_t near_at(const GridView& g, int32_t cx, int32_t cy, uint32_t c) {
  uint32_t s, e;
  for (int q = 0; q < 9; q++)
    if (near_cell(g, cx, cy, q, s, e)) {
      uint32_t len = e - s;
      if (c < len) return g.idx[s + c];
      c -= len;
    }
  return 0xFFFFFFFFu;
}

DEV GridScratch scratchA(const MgDrum& D) {
  GridScratch G;
  G.start = D.gA_start; G.idx = D.gA_idx; G.key = D.gA_key; G.meta = D.gA_meta; G.cursor = D.cursor;
  G.skey = D.gA_skey; G.sidx = D.gA_sidx; G.cap_cells = D.fcell_cap; G.pow2 = D.fpow2;
  return G;
}
DEV GridScratch scratchB(const MgDrum& D) {
  GridScratch G;
  G.start = D.gB_start; G.idx = D.gB_idx; G.key = D.gB_key; G.meta = D.gB_meta; G.cursor = D.cursor;
  G.skey = D.gB_skey; G.sidx = D.gB_sidx; G.cap_cells = D.fcell_cap; G.pow2 = D.fpow2;
  return G;
}
DEV GridScratch scratchC(const MgDrum& D) {
  GridScratch G;
  G.start = D.gC_start; G.idx = D.gC_idx; G.key = D.gC_key; G.meta = D.gC_meta; G.cursor = D.cursor;
  G.skey = D.gC_skey; G.sidx = D.gC_sidx; G.cap_cells = D.bcell_cap; G.pow2 = D.bpow2;
  return G;
}
DEV double sum_ke(const MgDrum& D, int n) {
  double s = 0.0;
  for (int i = tid(); i < n; i += nthr()) s += (double)vlen2(D.fv[i]);
  return 0.5 * (double)D.fmass * block_sum_d(s);
}
DEV bool fluid_active(const MgDrum& D, const MgSubstepIn& S) { return D.nf > 0 && S.dt > 0.0f; }

// ---------------------------------------------------------------------------------------------
// k_fluid_prep: sanitize + predict (pbf.rs:1200-1235), ball grid (:1214-1222), fluid grid and
// neighbour lists (:1238-1239, :262-288).
// ---------------------------------------------------------------------------------------------
KERNEL k_fluid_prep(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  MgScalars* sc = D.sc;
  MgStage* st = D.st;
  const int n = D.nf, nb = D.nb;
  for (int b = tid(); b < nb; b += nthr()) { D.imp[b] = v2(0.0f, 0.0f); D.ang[b] = 0.0f; }
  if (tid() == 0) {
    sc->visc_iters = 0; sc->density_iters = 0; sc->n_groups = 0; sc->n_links = 0; sc->links_in_range = 0;
    sc->mean_shear = 0.0f; sc->fluid_wall_work = 0.0f; sc->ke_noslip = sc->ke_drag = sc->ke_buoy = 0.0f;
    sc->ke_visc = sc->ke_clamp = 0.0f; sc->wall_imp_x = sc->wall_imp_y = sc->wall_ang = 0.0;
    for (int k = 0; k < 32; k++) sc->wall_bins[k] = 0.0f;
    st->ngroups = 0; st->nlinks = 0; st->in_range = 0;
  }
  if (!fluid_active(D, S)) return;
  const float dt = S.dt, h = D.h, br = D.br;
  float2* x = D.fx;
  float2* v = D.fv;
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = x[i], vi = v[i];
    if (!vfinite(xi) || !vfinite(vi)) { xi = v2(0.0f, 0.0f); vi = v2(0.0f, 0.0f); }
    D.x0[i] = xi;
    vi.y += GRAV * dt;
    xi = vadd(xi, vmul(vi, dt));
    x[i] = xi;
    v[i] = vi;
  }
  bsync();
  uint32_t* err = &sc->err;
  const float contact_radius = br + 0.05f * h;
  GridView gball = grid_build(D.bx, nb, fmaxf(fmaxf(2.0f * br, contact_radius), 1e-6f), scratchC(D), err);
  GridView g = grid_build(x, n, h, scratchA(D), err);
  // Neighbour lists in the reference's order: for a point in cell c, pair emission visits the
  // cells (-1,-1), (-1,0), (0,-1), c itself (cell order, skipping the point), (1,0), (1,1),
  // (0,1), (-1,1), and finally (1,-1).
  const float h2 = h * h;
  for (int pass = 0; pass < 2; pass++) {
    for (int i = tid(); i < n; i += nthr()) {
      float2 xi = x[i];
      uint32_t cnt = 0;
      uint32_t w = pass ? D.noff[i] : 0u;
      const uint32_t wend = pass ? D.noff[i + 1] : 0u;
      if (vfinite(xi) && g.ncells) {
        int32_t cx, cy;
        cell_of(xi, g.cell, cx, cy);
        for (int o = 0; o < 9; o++) {
          const int ox = o == 0 ? -1 : (o == 1 ? -1 : (o == 2 ? 0 : (o == 3 ? 0 : (o == 4 ? 1 : (o == 5 ? 1 : (o == 6 ? 0 : (o == 7 ? -1 : 1)))))));
          const int oy = o == 0 ? -1 : (o == 1 ? 0 : (o == 2 ? -1 : (o == 3 ? 0 : (o == 4 ? 0 : (o == 5 ? 1 : (o == 6 ? 1 : (o == 7 ? 1 : -1)))))));
          int32_t ax = sadd_i32(cx, ox), ay = sadd_i32(cy, oy);
          if (o != 3 && ax == cx && ay == cy) continue;
          uint32_t s, e;
          if (!grid_range(g, ax, ay, s, e)) continue;
          for (uint32_t a = s; a < e; a++) {
            uint32_t j = g.idx[a];
            if (j == (uint32_t)i) continue;
            if (vlen2(vsub(xi, x[j])) <= h2) {
              if (pass) {
                if (w < wend) D.nbr[w] = j;
                w++;
              } else {
                cnt++;
              }
            }
          }
        }
      }
      if (!pass) D.noff[i] = cnt;
    }
    bsync();
    if (!pass) {
      if (tid() == 0) D.noff[n] = 0;
      bsync();
      block_scan_excl(D.noff, n + 1);
      if (D.noff[n] > (uint32_t)D.e_cap) {
        // Capacity exceeded: report it and clip the lists so no later stage reads past nbr[].
        if (tid() == 0) atomicOr(err, MG_ERR_EDGES);
        bsync();
        for (int i = tid(); i <= n; i += nthr()) D.noff[i] = min(D.noff[i], (uint32_t)D.e_cap);
      }
      bsync();
    }
  }
  if (tid() == 0) { st->gball = gball; st->gA = g; }
}

// ---------------------------------------------------------------------------------------------
// k_density: density-constraint passes (pbf.rs:1257-1334), the gradient cache recomputed.
// ---------------------------------------------------------------------------------------------
KERNEL k_density(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  const int n = D.nf;
  const float h = D.h, rest = D.rest, mass = D.fmass;
  float2* x = D.fx;
  const float k6 = D.k_poly6, ks = D.k_spiky;
  const float m_over_rest = mass / rest;
  uint32_t base = max(D.pbf_iters, 1u), it_count = 0;
  for (;;) {
    float max_c = 0.0f;
    for (int i = tid(); i < n; i += nthr()) {
      float2 xi = x[i];
      float rho = mass * poly6_k(0.0f, h, k6);
      float2 gs = v2(0.0f, 0.0f);
      float sgs = 0.0f;
      for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
        float2 d = vsub(xi, x[D.nbr[k]]);
        float r2 = vlen2(d);
        rho += mass * poly6_k(r2, h, k6);
        float r = vlen(d);
        float2 g = vmul(spiky_grad_k(d, r, h, ks), m_over_rest);
        gs = vadd(gs, g);
        sgs += vlen2(g);
      }
      D.dens[i] = rho;
      float c = fmaxf(rho / rest - 1.0f, 0.0f);
      max_c = fmaxf(max_c, c);
      sgs += vlen2(gs);
      D.lam[i] = -c / (sgs + 200.0f);
    }
    max_c = block_max_f(max_c);
    for (int i = tid(); i < n; i += nthr()) {
      float2 xi = x[i];
      float li = D.lam[i];
      float2 s = v2(0.0f, 0.0f);
      for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
        uint32_t j = D.nbr[k];
        float2 d = vsub(xi, x[j]);
        float r = vlen(d);
        s = vadd(s, smul(li + D.lam[j], spiky_grad_k(d, r, h, ks)));
      }
      D.dp[i] = vmul(s, m_over_rest);
    }
    bsync();
    for (int i = tid(); i < n; i += nthr()) x[i] = vadd(x[i], D.dp[i]);
    bsync();
    it_count++;
    bool conv = max_c <= 0.02f;
    if (S.force_density) {
      if (it_count >= S.force_density) break;
    } else if ((it_count >= base && conv) || it_count >= base + 6) {
      break;
    }
  }
  if (tid() == 0) D.sc->density_iters = it_count;
}

// ---------------------------------------------------------------------------------------------
// k_couple: overlap push (pbf.rs:1361-1444), cohesion (:1464-1491), adhesion (:1515-1631).
// ---------------------------------------------------------------------------------------------
KERNEL k_couple(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  MgStage* st = D.st;
  uint32_t* err = &D.sc->err;
  const int n = D.nf, nb = D.nb;
  const float dt = S.dt, h = D.h, rest = D.rest, mass = D.fmass, br = D.br;
  float2* x = D.fx;
  float2* v = D.fv;
  const GridView gball = st->gball;
  const GridView gA = st->gA;
  __shared__ uint32_t s_nrec;
  if (tid() == 0) s_nrec = 0;
  bsync();
  // 3.5 overlap push: ball reactions become records (ball, fluid index), summed per ball in
  // ascending fluid index (the reference's accumulation order).
  if (nb > 0) {
    const float contact_radius = br + 0.05f * h;
    const float red = D.bmass > 0.0f ? (mass * D.bmass) / (mass + D.bmass) : mass;
    for (int i = tid(); i < n; i += nthr()) {
      float2 xi = x[i];
      const float2 xq = xi;  // the query uses the position before this particle's pushes
      float2 vi = v[i];
      float2 exc = v2(0.0f, 0.0f);
      if (vfinite(xq) && gball.ncells) {
        int32_t cx, cy;
        cell_of(xq, gball.cell, cx, cy);
        for (int q = 0; q < 9; q++) {
          uint32_t s, e;
          if (!near_cell(gball, cx, cy, q, s, e)) continue;
          for (uint32_t a = s; a < e; a++) {
            uint32_t b = gball.idx[a];
            float2 d = vsub(xi, D.bx[b]);
            float dist = vlen(d);
            if (isnan(dist) || dist >= contact_radius) continue;
            float2 nh = dist > 1e-9f ? vdiv(d, dist) : v2(1.0f, 0.0f);
            float pm = fminf(contact_radius - dist, 0.5f * br);
            xi = vadd(xi, smul(pm, nh));
            float2 lever = vmul(nh, br);
            float2 vs = vadd(D.bv[b], smul(D.bom[b], v2(-lever.y, lever.x)));
            float arrest = fmaxf(-vdot(vsub(vi, vs), nh), 0.0f);
            float full = mass * pm / dt;
            float im = fminf(full, red * arrest);
            exc = vadd(exc, smul((full - im) / mass, nh));
            float2 react = smul(-im, nh);
            uint32_t p = atomicAdd(&s_nrec, 1u);
            if (p < (uint32_t)D.rec_cap) {
              D.rec_key[p] = b; D.rec_ord[p] = (uint32_t)i;
              D.rec_v0[p] = react.x; D.rec_v1[p] = react.y; D.rec_v2[p] = cross2(lever, react);
            }
          }
        }
      }
      x[i] = xi;
      D.excess[i] = exc;
    }
    bsync();
    uint32_t nrec = s_nrec;
    if (nrec > (uint32_t)D.rec_cap) { if (tid() == 0) atomicOr(err, MG_ERR_RECS); nrec = D.rec_cap; }
    bsync();
    ordered_csr((int)nrec, D.rec_key, D.rec_ord, nb, D.rec_seg, D.rec_perm, D.cursor);
    for (int b = tid(); b < nb; b += nthr()) {
      float2 im = D.imp[b];
      float an = D.ang[b];
      for (uint32_t q = D.rec_seg[b]; q < D.rec_seg[b + 1]; q++) {
        uint32_t r = D.rec_perm[q];
        im = vadd(im, v2(D.rec_v0[r], D.rec_v1[r]));
        an += D.rec_v2[r];
      }
      D.imp[b] = im;
      D.ang[b] = an;
    }
    bsync();
  } else {
    for (int i = tid(); i < n; i += nthr()) D.excess[i] = v2(0.0f, 0.0f);
    bsync();
  }
  // 3.6a cohesion
  const float cacc = (fmaxf(D.surface_tension, 0.0f) / 0.072f) * 2.0f * 9.81f;
  for (int i = tid(); i < n; i += nthr()) {
    float2 pull = v2(0.0f, 0.0f);
    if (cacc > 0.0f) {
      float2 xi = x[i];
      for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
        float2 d = vsub(xi, x[D.nbr[k]]);
        float r = vlen(d);
        if (r <= 1e-9f) continue;
        float w = mass * cohesion_kernel_k(r, h, D.k_coh, D.h6_64) / rest;
        pull = vsub(pull, smul(w, vdiv(d, r)));
      }
      D.coh[i] = vmul(smul(cacc, pull), dt);
    } else {
      D.coh[i] = v2(0.0f, 0.0f);
    }
    D.adh[i] = v2(0.0f, 0.0f);
  }
  bsync();
  // 3.6b adhesion: one thread per ball merges its boundary points' contributions in
  // first-appearance order; the fluid kicks are summed per particle in ball order.
  const float wett = fclamp(D.wettability, 0.0f, 1.0f);
  const float ha = D.h_adh;
  if (wett > 0.0f && nb > 0 && br > 0.0f && 0.5f * h > 0.0f && ha > 0.0f) {
    const float acc = wett * 2.0f * 9.81f;
    const float range2 = ha * ha * (1.0f + 1e-4f);
    const float inner2 = 0.25f * ha * ha * (1.0f - 1e-4f);
    if (tid() == 0) s_nrec = 0;
    bsync();
    for (int b = tid(); b < nb; b += nthr()) {
      float2 xb = D.bx[b];
      if (!vfinite(xb)) continue;
      uint32_t* ij = D.it_j + (size_t)b * D.item_cap;
      float2* iv = D.it_v + (size_t)b * D.item_cap;
      float* iw = D.it_w + (size_t)b * D.item_cap;
      uint32_t ni = 0;
      bool over = false;
      for (uint32_t k = 0; k < D.bp_cnt && k < MG_MAX_BP; k++) {
        float2 bp = vadd(xb, v2(D.bp_x[k], D.bp_y[k]));
        if (!vfinite(bp) || !gA.ncells) continue;
        int32_t cx, cy;
        cell_of(bp, gA.cell, cx, cy);
        for (int q = 0; q < 9; q++) {
          uint32_t s, e;
          if (!near_cell(gA, cx, cy, q, s, e)) continue;
          for (uint32_t a = s; a < e; a++) {
            uint32_t j = gA.idx[a];
            float2 d = vsub(bp, x[j]);
            float r2 = vlen2(d);
            if (r2 >= range2 || r2 <= inner2) continue;
            float r = sqrtf(r2);
            if (!isfinite(r) || r <= 0.5f * ha || r >= ha) continue;
            float w = mass * adhesion_kernel_k(r, ha, D.k_adh) / rest;
            if (w <= 0.0f) continue;
            float2 nh = vdiv(d, r);
            uint32_t sl = 0;
            while (sl < ni && ij[sl] != j) sl++;
            if (sl == ni) {
              if (ni < (uint32_t)D.item_cap) { ij[ni] = j; iv[ni] = smul(w, nh); iw[ni] = w; ni++; }
              else over = true;
            } else {
              iv[sl] = vadd(iv[sl], smul(w, nh));
              iw[sl] += w;
            }
          }
        }
      }
      if (over) atomicOr(err, MG_ERR_ITEMS);
      if (ni == 0) continue;
      float wsum = 0.0f;
      float2 pds = v2(0.0f, 0.0f);
      for (uint32_t sl = 0; sl < ni; sl++) { pds = vadd(pds, iv[sl]); wsum += iw[sl]; }
      if (wsum <= 1e-9f) continue;
      float pl = vlen(pds);
      if (pl <= 1e-9f) continue;
      float2 bdir = vdiv(vneg(pds), pl);
      float cov = fminf(wsum, 1.0f);
      float mc = mass * wsum;
      float att = mc / (D.bmass + mc);
      float2 dvb = vmul(bdir, acc * cov * dt * att);
      D.imp[b] = vadd(D.imp[b], smul(D.bmass, dvb));
      float2 lever = vmul(bdir, br);
      D.ang[b] += D.bmass * cross2(lever, dvb);
      for (uint32_t sl = 0; sl < ni; sl++) {
        float frac = iw[sl] / wsum;
        float2 dvi = smul(-(D.bmass * frac / mass), dvb);
        uint32_t p = atomicAdd(&s_nrec, 1u);
        if (p < (uint32_t)D.rec_cap) {
          D.rec_key[p] = ij[sl]; D.rec_ord[p] = (uint32_t)b; D.rec_v0[p] = dvi.x; D.rec_v1[p] = dvi.y;
        }
      }
    }
    bsync();
    uint32_t nrec = s_nrec;
    if (nrec > (uint32_t)D.rec_cap) { if (tid() == 0) atomicOr(err, MG_ERR_RECS); nrec = D.rec_cap; }
    bsync();
    ordered_csr((int)nrec, D.rec_key, D.rec_ord, n, D.rec_seg, D.rec_perm, D.cursor);
    for (int i = tid(); i < n; i += nthr()) {
      float2 a = v2(0.0f, 0.0f);
      for (uint32_t q = D.rec_seg[i]; q < D.rec_seg[i + 1]; q++) {
        uint32_t r = D.rec_perm[q];
        a = vadd(a, v2(D.rec_v0[r], D.rec_v1[r]));
      }
      D.adh[i] = a;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// k_wall: wall projection, velocity reconstruction, no-slip blend (pbf.rs:1635-1700).
// ---------------------------------------------------------------------------------------------
KERNEL k_wall(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  MgScalars* sc = D.sc;
  const int n = D.nf;
  const float dt = S.dt, mass = D.fmass;
  float2* x = D.fx;
  float2* v = D.fv;
  __shared__ double s_bins[32];
  if (tid() < 32) s_bins[tid()] = 0.0;
  bsync();
  double wix = 0.0, wiy = 0.0, wan = 0.0;
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = x[i];
    float2 nrm;
    float d = drum_sdf_world(D, xi, S.cn, S.sn, S.cp, S.sp, nrm);
    uint32_t touched = 0;
    if (d < 0.0f) {
      xi = vadd(xi, smul(-d, nrm));
      touched = 1;
      float2 push = vdiv(smul(mass, smul(-d, nrm)), dt);
      wix += (double)push.x; wiy += (double)push.y;
      wan += (double)cross2(xi, push);
      float a = atan2f(xi.y, xi.x);
      uint32_t bin = sat_u32(((a + 3.14159265f) / 6.28318531f) * 32.0f);
      atomicAdd(&s_bins[min(bin, 31u)], (double)(mass * -d / dt));
    }
    x[i] = xi;
    D.touched[i] = touched;
    float2 vi = vdiv(vsub(xi, D.x0[i]), dt);
    vi = vsub(vi, D.excess[i]);
    vi = vadd(vi, D.coh[i]);
    vi = vadd(vi, D.adh[i]);
    v[i] = vi;
  }
  bsync();
  const double ke_before = sum_ke(D, n);
  const float beta = fclamp(D.wall_no_slip, 0.0f, 1.0f);
  double ww = 0.0;
  if (beta > 0.0f) {
    for (int i = tid(); i < n; i += nthr()) {
      if (!D.touched[i]) continue;
      float2 xi = x[i];
      float2 vw = wall_velocity(D.omega, xi);
      float2 vo = v[i];
      float2 vn = vadd(smul(1.0f - beta, vo), smul(beta, vw));
      v[i] = vn;
      float2 dv = vsub(vn, vo);
      ww += (double)(mass * vdot(dv, vw));
      wix += (double)(mass * dv.x); wiy += (double)(mass * dv.y);
      wan += (double)cross2(xi, smul(mass, dv));
    }
  }
  bsync();
  ww = block_sum_d(ww);
  wix = block_sum_d(wix);
  wiy = block_sum_d(wiy);
  wan = block_sum_d(wan);
  const double ke_mark = sum_ke(D, n);
  if (tid() == 0) {
    sc->ke_noslip = (float)(ke_mark - ke_before);
    sc->fluid_wall_work = (float)ww;
    sc->wall_imp_x = wix; sc->wall_imp_y = wiy; sc->wall_ang = wan;
    for (int k = 0; k < 32; k++) sc->wall_bins[k] = (float)s_bins[k];
    D.st->ke_mark = ke_mark;
  }
}

// ---------------------------------------------------------------------------------------------
// k_drag: form drag (pbf.rs:1777-2013), Gauss-Seidel over balls in index order by wave 0.
// ---------------------------------------------------------------------------------------------
KERNEL k_drag(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  MgScalars* sc = D.sc;
  uint32_t* err = &sc->err;
  const int n = D.nf, nb = D.nb;
  const float dt = S.dt, rest = D.rest, mass = D.fmass, br = D.br;
  float2* x = D.fx;
  float2* v = D.fv;
  const float mu = fmaxf(D.viscosity, 0.0f);
  const float beta_b = fclamp(D.ball_no_slip, 0.0f, 1.0f);
  if (nb > 0 && br > 0.0f && D.btrue > 0.0f && mu > 1e-6f && beta_b > 0.0f) {
    const float hc = D.h_c;
    GridView gB = grid_build(x, n, fmaxf(hc, 1e-6f), scratchB(D), err);
    if (tid() < 64) {
      const int lane = lane_id();
      const float rot_cap = 0.2f * D.v_max;
      const float rho_ball = D.bmass / (3.14159265f * br * br);
      const float mass_true = rho_ball * 3.14159265f * D.btrue * D.btrue;
      const float ib = D.binertia;
      const float hc2 = hc * hc;
      const uint32_t cap = (uint32_t)D.item_cap;
      uint32_t* ij = D.it_j;
      float2* iv = D.it_v;  // r_i
      float* iw = D.it_w;   // phi
      for (int b = 0; b < nb; b++) {
        float2 xb = D.bx[b];
        if (!vfinite(xb) || !gB.ncells) continue;
        int32_t cx, cy;
        cell_of(xb, gB.cell, cx, cy);
        const uint32_t total = near_total(gB, cx, cy);
        if (total == 0) continue;
        float m_ent = 0.0f, i_ent = 0.0f, l_sum = 0.0f;
        float2 r_sum = v2(0.0f, 0.0f), v_sum = v2(0.0f, 0.0f);
        uint32_t ni = 0;
        for (uint32_t base = 0; base < total; base += 64) {
          uint32_t c = base + lane;
          bool ok = false;
          uint32_t j = 0;
          float2 ri = v2(0.0f, 0.0f), vj = v2(0.0f, 0.0f);
          float r2 = 0.0f, phi = 0.0f;
          if (c < total) {
            j = near_at(gB, cx, cy, c);
            if (j < (uint32_t)n) {
              ri = vsub(x[j], xb);
              r2 = vlen2(ri);
              if (isfinite(r2) && r2 < hc2) {
                phi = poly6_k(r2, hc, D.k_poly6_hc) / D.w_hc0;
                ok = phi > 0.0f;
                vj = v[j];
              }
            }
          }
          uint64_t m = __ballot(ok);
          uint32_t before = __popcll(m & ((1ull << lane) - 1ull));
          if (ok && ni + before < cap) { ij[ni + before] = j; iv[ni + before] = ri; iw[ni + before] = phi; }
          ni += __popcll(m);
          while (m) {
            int s = next_lane(m);
            float w = mass * rlf(phi, s);
            float2 r_i = v2(rlf(ri.x, s), rlf(ri.y, s));
            float2 v_j = v2(rlf(vj.x, s), rlf(vj.y, s));
            m_ent += w;
            r_sum = vadd(r_sum, smul(w, r_i));
            v_sum = vadd(v_sum, smul(w, v_j));
            i_ent += w * rlf(r2, s);
            l_sum += w * cross2(r_i, v_j);
          }
        }
        if (ni > cap) { if (lane == 0) atomicOr(err, MG_ERR_ITEMS); ni = cap; }
        if (m_ent <= 1e-12f || ni == 0) continue;
        __threadfence_block();
        float2 r_bar = vdiv(r_sum, m_ent), v_bar = vdiv(v_sum, m_ent);
        float2 v_rel = vsub(v_bar, D.bv[b]);
        float sp = vlen(v_rel);
        float inv_tau_form = (sp > 1e-6f && rest > 0.0f) ? 0.5f * rest * 1.0f * (2.0f * D.btrue) * sp / mass_true : 0.0f;
        float bt = beta_b * (1.0f - expf(-dt * (0.0f + inv_tau_form)));
        float a_lin = bt * m_ent / (D.bmass + m_ent);
        float2 dvb = smul(a_lin, v_rel);
        float dwb = 0.0f, c_rot = 0.0f;
        float2 p_bar = v2(0.0f, 0.0f);
        if (i_ent > 1e-12f) {
          float om_bar = l_sum / i_ent;
          float a_rot = bt * i_ent / (ib + i_ent);
          float cand = a_rot * (om_bar - D.bom[b]);
          p_bar = v2(-r_bar.y, r_bar.x);
          float s = 0.0f, s_scale = 0.0f;
          for (uint32_t base = 0; base < ni; base += 64) {
            uint32_t q = base + lane;
            float ts = 0.0f, tss = 0.0f;
            if (q < ni) {
              float2 r_i = iv[q];
              float ph = iw[q];
              float2 pi_ = v2(-r_i.y, r_i.x);
              float phi2 = ph * ph;
              ts = mass * phi2 * cross2(r_i, vsub(pi_, p_bar));
              tss = mass * phi2 * vlen2(r_i);
            }
            uint32_t cnt = min(64u, ni - base);
            for (uint32_t l = 0; l < cnt; l++) { s += rlf(ts, (int)l); s_scale += rlf(tss, (int)l); }
          }
          if (fabsf(s) > 0.1f * s_scale) {
            float raw = -ib * cand / s;
            float ml = 2.0f * hc;
            float capr = ml > 1e-9f ? rot_cap / ml : 0.0f;
            c_rot = fclamp(raw, -capr, capr);
            dwb = -c_rot * s / ib;
          }
        }
        for (uint32_t q = lane; q < ni; q += 64) {
          float2 r_i = iv[q];
          float ph = iw[q];
          float2 dvi = vmul(smul(-(D.bmass / m_ent), dvb), ph);
          if (c_rot != 0.0f) {
            float2 pi_ = v2(-r_i.y, r_i.x);
            dvi = vadd(dvi, smul(c_rot * ph, vsub(pi_, p_bar)));
          }
          uint32_t j = ij[q];
          if (j < (uint32_t)n) v[j] = vadd(v[j], dvi);
        }
        if (lane == 0) {
          D.imp[b] = vadd(D.imp[b], smul(D.bmass, dvb));
          D.ang[b] += ib * dwb + D.bmass * cross2(r_bar, dvb);
        }
        __threadfence_block();
      }
    }
    bsync();
  }
  const double ke_after = sum_ke(D, n);
  if (tid() == 0) {
    sc->ke_drag = (float)(ke_after - D.st->ke_mark);
    D.st->ke_mark = ke_after;
  }
}

// ---------------------------------------------------------------------------------------------
// k_buoy: buoyancy (pbf.rs:2035-2122; step-2 grid, current positions, balls in index order),
// density recompute (:2136-2144), effective viscosity and Morris weights (:2147-2154, :312-344).
// ---------------------------------------------------------------------------------------------
KERNEL k_buoy(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  MgScalars* sc = D.sc;
  uint32_t* err = &sc->err;
  const int n = D.nf, nb = D.nb;
  const float dt = S.dt, h = D.h, rest = D.rest, mass = D.fmass, br = D.br;
  float2* x = D.fx;
  float2* v = D.fv;
  const GridView gA = D.st->gA;
  if (nb > 0 && br > 0.0f) {
    if (tid() < 64) {
      const int lane = lane_id();
      const float area = 3.14159265f * br * br;
      const float om_abs = fabsf(D.omega);
      const uint32_t cap = (uint32_t)D.item_cap;
      uint32_t* ij = D.it_j;
      float* iw = D.it_w;
      for (int b = 0; b < nb; b++) {
        float2 xb = D.bx[b];
        if (!vfinite(xb) || !gA.ncells) continue;
        int32_t cx, cy;
        cell_of(xb, gA.cell, cx, cy);
        const uint32_t total = near_total(gA, cx, cy);
        if (total == 0) continue;
        float rho_l = 0.0f, on = 0.0f, od = 0.0f;
        uint32_t ni = 0;
        for (uint32_t base = 0; base < total; base += 64) {
          uint32_t c = base + lane;
          bool ok = false, dok = false;
          uint32_t j = 0;
          float w = 0.0f, tn = 0.0f;
          if (c < total) {
            j = near_at(gA, cx, cy, c);
            if (j < (uint32_t)n) {
              float r2 = vlen2(vsub(xb, x[j]));
              w = mass * poly6_k(r2, h, D.k_poly6);
              ok = w > 0.0f;
              if (ok) {
                float2 dj = x[j];
                float d2 = vlen2(dj);
                if (!(d2 <= 1e-12f)) { dok = true; tn = w * cross2(dj, v[j]) / d2; }
              }
            }
          }
          uint64_t m = __ballot(ok);
          uint32_t before = __popcll(m & ((1ull << lane) - 1ull));
          if (ok && ni + before < cap) { ij[ni + before] = j; iw[ni + before] = w; }
          ni += __popcll(m);
          uint64_t m2 = m;
          while (m2) { int s = next_lane(m2); rho_l += rlf(w, s); }
          uint64_t md = __ballot(dok);
          while (md) { int s = next_lane(md); on += rlf(tn, s); od += rlf(w, s); }
        }
        if (ni > cap) { if (lane == 0) atomicOr(err, MG_ERR_ITEMS); ni = cap; }
        if (rho_l <= 0.0f) continue;
        __threadfence_block();
        float rho_eff = fminf(rho_l, rest);
        float om_l = od > 0.0f ? fclamp(on / od, -om_abs, om_abs) : 0.0f;
        float2 ge = vadd(v2(0.0f, GRAV), smul(om_l * om_l, xb));
        float2 impb = vmul(smul((-rho_eff) * area, ge), dt);
        for (uint32_t q = lane; q < ni; q += 64) {
          float frac = iw[q] / rho_l;
          float2 dv = vdiv(vneg(vmul(impb, frac)), mass);
          uint32_t j = ij[q];
          if (j < (uint32_t)n) v[j] = vadd(v[j], dv);
        }
        if (lane == 0) D.imp[b] = vadd(D.imp[b], impb);
        __threadfence_block();
      }
    }
    bsync();
  }
  const double ke_after = sum_ke(D, n);
  if (tid() == 0) {
    sc->ke_buoy = (float)(ke_after - D.st->ke_mark);
    D.st->ke_mark = ke_after;
  }
  // 7a density at the final positions
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = x[i];
    float rho = mass * poly6_k(0.0f, h, D.k_poly6);
    for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) rho += mass * poly6_k(vlen2(vsub(xi, x[D.nbr[k]])), h, D.k_poly6);
    D.dens[i] = rho;
  }
  bsync();
  const float mu = fmaxf(D.viscosity, 0.0f);
  if (!(mu > 0.0f)) return;
  // 7b effective viscosity (Bingham, pbf.rs:821-868) and Morris weights (pbf.rs:312-344)
  const bool bingham = D.rheology == 1 && D.yield_stress > 0.0f;
  for (int i = tid(); i < n; i += nthr()) {
    float me = mu;
    if (bingham) {
      float2 xi = x[i], vi = v[i];
      float gxx = 0.0f, gxy = 0.0f, gyx = 0.0f, gyy = 0.0f;
      for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
        uint32_t j = D.nbr[k];
        float2 d = vsub(xi, x[j]);
        float r = vlen(d);
        float2 g = spiky_grad_k(d, r, h, D.k_spiky);
        float coeff = mass / fmaxf(D.dens[j], 1e-6f);
        float2 dv = vsub(v[j], vi);
        gxx += coeff * dv.x * g.x; gxy += coeff * dv.x * g.y;
        gyx += coeff * dv.y * g.x; gyy += coeff * dv.y * g.y;
      }
      float dxy = 0.5f * (gxy + gyx);
      float rate = sqrtf(2.0f * (gxx * gxx + gyy * gyy + 2.0f * dxy * dxy));
      float tau_y = D.yield_stress;
      float g0 = tau_y / (100.0f * mu);
      float g = isfinite(rate) ? rate : 0.0f;
      float extra = g < 1e-4f * g0 ? tau_y / g0 : tau_y * (1.0f - expf(-g / g0)) / g;
      me = mu + extra;
    }
    D.mu_eff[i] = me;
  }
  bsync();
  const float eta2 = 0.01f * h * h;
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = x[i];
    float rho_i = fmaxf(D.dens[i], 1e-6f);
    float mui = D.mu_eff[i];
    for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
      uint32_t j = D.nbr[k];
      float rho_j = fmaxf(D.dens[j], 1e-6f);
      float2 d = vsub(xi, x[j]);
      float r = vlen(d);
      float2 g = spiky_grad_k(d, r, h, D.k_spiky);
      float num = (mass * (mui + D.mu_eff[j]) / (rho_i * rho_j)) * fabsf(vdot(d, g));
      D.W[k] = num / (vlen2(d) + eta2);
      D.D[k] = r > 1e-9f ? vdiv(d, r) : v2(1.0f, 0.0f);
    }
    D.ffac[i] = 0.5f * (mass * 2.0f * mui / (rho_i * rest));
  }
}

// ---------------------------------------------------------------------------------------------
// k_links: ball boundary groups (pbf.rs:2159-2232) and squeeze links (pbf.rs:544-625).
// ---------------------------------------------------------------------------------------------
DEV double resolved_ratio(double hcl) {  // lubrication.rs:95-115
  if (hcl >= 8.0) return 1.0;
  double hh = fmax(hcl, 0.0);
  if (hh <= 0.5) return 0.0 + (hh - 0.0) / (0.5 - 0.0) * (0.06 - 0.0);
  if (hh <= 1.0) return 0.06 + (hh - 0.5) / (1.0 - 0.5) * (0.25 - 0.06);
  if (hh <= 2.0) return 0.25 + (hh - 1.0) / (2.0 - 1.0) * (0.90 - 0.25);
  if (hh <= 4.0) return 0.90 + (hh - 2.0) / (4.0 - 2.0) * (0.99 - 0.90);
  if (hh <= 8.0) return 0.99 + (hh - 4.0) / (8.0 - 4.0) * (1.0 - 0.99);
  return 1.0;
}
DEV float link_coefficient(const double* film, double a, double dxd, float gap, float c_max) {
  double wgt = 1.0 - resolved_ratio((double)gap / dxd);
  double hh = fmax((double)gap, 0.02 * a);
  double eps = hh / film[0];
  double cf = film[3] / (hh * sqrt(hh)) * (1.0 + film[1] * eps + film[2] * eps * eps);
  double c = wgt * cf;
  return fminf(1.0f * (float)c, c_max);
}

KERNEL k_links(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  const float mu = fmaxf(D.viscosity, 0.0f);
  if (!(mu > 0.0f)) return;
  MgScalars* sc = D.sc;
  MgStage* st = D.st;
  uint32_t* err = &sc->err;
  const int n = D.nf, nb = D.nb;
  const float dt = S.dt, h = D.h, br = D.br;
  float2* x = D.fx;
  const float eta2 = 0.01f * h * h;
  uint32_t ngroups = 0;
  if (nb > 0 && br > 0.0f && 0.5f * h > 0.0f) {
    GridView gnow = grid_build(x, n, h, scratchB(D), err);
    const float sn = D.k_spiky;
    const float range2 = h * h * (1.0f + 1e-4f);
    for (int b = tid(); b < nb; b += nthr()) {
      uint32_t ni = 0;
      float2 xb = D.bx[b];
      uint32_t* ij = D.it_j + (size_t)b * D.item_cap;
      float2* iv = D.it_v + (size_t)b * D.item_cap;
      float* iw = D.it_w + (size_t)b * D.item_cap;
      float* iw2 = D.it_w2 + (size_t)b * D.item_cap;
      bool over = false;
      if (vfinite(xb) && gnow.ncells) {
        for (uint32_t k = 0; k < D.bp_cnt && k < MG_MAX_BP; k++) {
          float2 bp = vadd(xb, v2(D.bp_x[k], D.bp_y[k]));
          if (!vfinite(bp)) continue;
          int32_t cx, cy;
          cell_of(bp, gnow.cell, cx, cy);
          for (int q = 0; q < 9; q++) {
            uint32_t s, e;
            if (!near_cell(gnow, cx, cy, q, s, e)) continue;
            for (uint32_t a = s; a < e; a++) {
              uint32_t j = gnow.idx[a];
              float2 d = vsub(x[j], bp);
              float r2 = vlen2(d);
              if (r2 >= range2) continue;
              float r = sqrtf(r2);
              if (!isfinite(r) || r >= h || r <= 1e-9f) continue;
              float da = fabsf(sn * (h - r) * (h - r) * r);
              float c = D.ffac[j] * da / (r2 + eta2);
              if (!(c > 0.0f)) continue;
              float2 lever = vsub(bp, xb);
              float2 lc = smul(c, lever);
              float l2 = c * vlen2(lever);
              uint32_t sl = 0;
              while (sl < ni && ij[sl] != j) sl++;
              if (sl == ni) {
                if (ni < (uint32_t)D.item_cap) { ij[ni] = j; iv[ni] = lc; iw[ni] = c; iw2[ni] = l2; ni++; }
                else over = true;
              } else {
                iw[sl] += c; iv[sl] = vadd(iv[sl], lc); iw2[sl] += l2;
              }
            }
          }
        }
      }
      if (over) atomicOr(err, MG_ERR_ITEMS);
      D.it_n[b] = ni;
    }
    bsync();
    for (int b = tid(); b < nb; b += nthr()) D.g_bstart[b] = D.it_n[b];
    if (tid() == 0) D.g_bstart[nb] = 0;
    bsync();
    block_scan_excl(D.g_bstart, nb + 1);
    ngroups = D.g_bstart[nb];
    if (ngroups > (uint32_t)D.grp_cap) { if (tid() == 0) atomicOr(err, MG_ERR_GROUPS); ngroups = 0; }
    bsync();
    if (ngroups) {
      for (int b = tid(); b < nb; b += nthr()) {
        uint32_t o = D.g_bstart[b], ni = D.it_n[b];
        size_t base = (size_t)b * D.item_cap;
        for (uint32_t sl = 0; sl < ni; sl++) {
          D.g_fluid[o + sl] = D.it_j[base + sl];
          D.g_ball[o + sl] = (uint32_t)b;
          D.g_c[o + sl] = D.it_w[base + sl];
          D.g_lc[o + sl] = D.it_v[base + sl];
          D.g_l2[o + sl] = D.it_w2[base + sl];
        }
      }
      bsync();
      for (int gi = tid(); gi < (int)ngroups; gi += nthr()) D.rec_ord[gi] = (uint32_t)gi;
      bsync();
      // per-fluid group lists in group order (the CG apply's accumulation order)
      ordered_csr((int)ngroups, D.g_fluid, D.rec_ord, n, D.gf_start, D.gf_perm, D.cursor);
    }
  }
  // squeeze links, active set from the velocities after the earlier impulses
  for (int b = tid(); b < nb; b += nthr()) D.bv0[b] = vadd(D.bv[b], vdiv(D.imp[b], D.bmass));
  bsync();
  uint32_t nlinks = 0, in_range = 0;
  if (nb > 0 && br > 0.0f) {
    const float reach = D.reach;
    const float c_max = 8.0f * D.bmass / fmaxf(dt, 1e-12f);
    GridView g2 = grid_build(D.bx, nb, 2.0f * br + reach, scratchC(D), err);
    const float rng2 = (2.0f * br + reach) * (2.0f * br + reach) * (1.0f + 1e-4f);
    const double a = (double)br, dxd = D.dxd;
    for (int pass = 0; pass < 2; pass++) {
      uint32_t my_inr = 0;
      for (int i = tid(); i < nb; i += nthr()) {
        float2 xi = D.bx[i];
        uint32_t cnt = 0;
        uint32_t wpos = pass ? D.l_bstart[i] : 0u;
        const uint32_t wend = pass ? D.l_bstart[i + 1] : 0u;
        if (vfinite(xi)) {
          if (g2.ncells) {
            int32_t cx, cy;
            cell_of(xi, g2.cell, cx, cy);
            for (int q = 0; q < 9; q++) {
              uint32_t s, e;
              if (!near_cell(g2, cx, cy, q, s, e)) continue;
              for (uint32_t aa = s; aa < e; aa++) {
                uint32_t j = g2.idx[aa];
                if ((int)j <= i) continue;
                float2 ev = vsub(D.bx[j], xi);
                if (vlen2(ev) >= rng2 || !vfinite(ev)) continue;
                float dist = vlen(ev);
                float gap = dist - 2.0f * br;
                if (gap >= reach || dist < 1e-9f) continue;
                my_inr++;
                float2 nn = vdiv(ev, dist);
                if (vdot(nn, vsub(D.bv0[i], D.bv0[j])) <= 0.0f) continue;
                if (pass) {
                  if (wpos < wend && wpos < (uint32_t)D.link_cap) {
                    D.l_ball[wpos] = (uint32_t)i; D.l_other[wpos] = j; D.l_n[wpos] = nn;
                    D.l_c[wpos] = link_coefficient(D.sq_pair, a, dxd, gap, c_max);
                  }
                  wpos++;
                } else {
                  cnt++;
                }
              }
            }
          }
          float dist = vlen(xi);
          float gap = D.R - br - dist;
          if (gap < reach && dist > 1e-9f) {
            my_inr++;
            float2 nn = vdiv(xi, dist);
            if (!(vdot(nn, D.bv0[i]) <= 0.0f)) {
              if (pass) {
                if (wpos < wend && wpos < (uint32_t)D.link_cap) {
                  D.l_ball[wpos] = (uint32_t)i; D.l_other[wpos] = MG_NO_SLOT; D.l_n[wpos] = nn;
                  D.l_c[wpos] = link_coefficient(D.sq_wall, a, dxd, gap, c_max);
                }
                wpos++;
              } else {
                cnt++;
              }
            }
          }
        }
        if (!pass) D.l_bstart[i] = cnt;
      }
      bsync();
      if (!pass) {
        in_range = block_sum_u(my_inr);
        if (tid() == 0) D.l_bstart[nb] = 0;
        bsync();
        block_scan_excl(D.l_bstart, nb + 1);
        nlinks = D.l_bstart[nb];
        if (nlinks > (uint32_t)D.link_cap) { if (tid() == 0) atomicOr(err, MG_ERR_LINKS); }
        bsync();
      }
    }
    if (nlinks > (uint32_t)D.link_cap) nlinks = 0;
    if (nlinks) {
      // per-ball incidence lists (ball as `ball` or as `other`), in link order
      for (int l = tid(); l < (int)nlinks; l += nthr()) {
        D.rec_key[2 * l] = D.l_ball[l];
        D.rec_ord[2 * l] = 2u * (uint32_t)l;
        uint32_t o = D.l_other[l];
        D.rec_key[2 * l + 1] = o == MG_NO_SLOT ? (uint32_t)nb : o;  // wall -> dummy key nb
        D.rec_ord[2 * l + 1] = 2u * (uint32_t)l + 1u;
      }
      bsync();
      ordered_csr((int)(2 * nlinks), D.rec_key, D.rec_ord, nb + 1, D.li_start, D.li_perm, D.cursor);
    }
  }
  if (tid() == 0) {
    sc->n_groups = ngroups; sc->n_links = nlinks; sc->links_in_range = in_range;
    st->ngroups = ngroups; st->nlinks = nlinks; st->in_range = in_range;
  }
}

// ---------------------------------------------------------------------------------------------
// k_visc: implicit viscosity (pbf.rs:423-480 fluid only, pbf.rs:641-791 joint with the balls).
// Dot products and residual norms accumulate in f64 (fixed reduction order).
// ---------------------------------------------------------------------------------------------
struct CgOp {
  int n, nb, len;
  bool joint;
  uint32_t ngroups, nlinks;
  float dt, mass, k_dm;
};
// y = A x  (joint: the packed operator of pbf.rs:682-731; fluid only: I + dt L, pbf.rs:395-407)
DEV void cg_apply(const MgDrum& D, const CgOp& op, const float2* xx, float2* out) {
  const int n = op.n, nb = op.nb;
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = xx[i];
    float2 s = v2(0.0f, 0.0f);
    for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
      float2 e = D.D[k];
      s = vadd(s, smul(4.0f * D.W[k] * vdot(e, vsub(xi, xx[D.nbr[k]])), e));
    }
    if (!op.joint) {
      out[i] = vadd(xi, smul(op.dt, s));
    } else {
      float2 o = smul(op.mass, vadd(xi, smul(op.dt, s)));
      if (op.ngroups) {
        for (uint32_t q = D.gf_start[i]; q < D.gf_start[i + 1]; q++) {
          uint32_t g = D.gf_perm[q];
          uint32_t bi = D.g_ball[g];
          float2 vb = xx[n + bi];
          float w = xx[n + nb + bi].x;
          float2 rel = vsub(xi, vb);
          o = vadd(o, smul(op.k_dm, vsub(smul(D.g_c[g], rel), smul(w, vperp(D.g_lc[g])))));
        }
      }
      out[i] = o;
    }
  }
  if (op.joint) {
    for (int b = tid(); b < nb; b += nthr()) {
      float2 vb = xx[n + b];
      float w = xx[n + nb + b].x;
      float2 ob = smul(D.bmass, vb);
      float ow = D.binertia * w;
      if (op.ngroups) {
        for (uint32_t g = D.g_bstart[b]; g < D.g_bstart[b + 1]; g++) {
          float2 rel = vsub(xx[D.g_fluid[g]], vb);
          float2 lc = D.g_lc[g];
          float2 d = smul(op.k_dm, vsub(smul(D.g_c[g], rel), smul(w, vperp(lc))));
          ob = vsub(ob, d);
          ow -= op.k_dm * (cross2(lc, rel) - w * D.g_l2[g]);
        }
      }
      if (op.nlinks) {
        for (uint32_t q = D.li_start[b]; q < D.li_start[b + 1]; q++) {
          uint32_t rr = D.li_perm[q];
          uint32_t l = D.rec_ord[rr] >> 1;
          bool as_other = (D.rec_ord[rr] & 1u) != 0u;
          uint32_t li = D.l_ball[l], lo = D.l_other[l];
          float2 rel = xx[n + li];
          if (lo != MG_NO_SLOT) rel = vsub(rel, xx[n + lo]);
          float2 nn = D.l_n[l];
          float2 d = smul(op.dt * D.l_c[l] * vdot(nn, rel), nn);
          ob = as_other ? vsub(ob, d) : vadd(ob, d);
        }
      }
      out[n + b] = ob;
      out[n + nb + b] = v2(ow, 0.0f);
    }
  }
  bsync();
}
DEV double cg_dot(int len, const float2* a, const float2* b) {
  double s = 0.0;
  for (int i = tid(); i < len; i += nthr()) s += (double)a[i].x * (double)b[i].x + (double)a[i].y * (double)b[i].y;
  return block_sum_d(s);
}
DEV double cg_res_norm(int len, const float* inv, const float2* r) {
  double s = 0.0;
  for (int i = tid(); i < len; i += nthr()) s += (double)inv[i] * ((double)r[i].x * (double)r[i].x + (double)r[i].y * (double)r[i].y);
  return sqrt(block_sum_d(s));
}

KERNEL k_visc(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  const float mu = fmaxf(D.viscosity, 0.0f);
  if (!(mu > 0.0f)) return;
  MgScalars* sc = D.sc;
  const MgStage* st = D.st;
  CgOp op;
  op.n = D.nf; op.nb = D.nb;
  op.ngroups = st->ngroups; op.nlinks = st->nlinks;
  op.joint = !(op.ngroups == 0 && st->in_range == 0);
  op.len = op.joint ? op.n + 2 * op.nb : op.n;
  op.dt = S.dt; op.mass = D.fmass; op.k_dm = S.dt * D.fmass;
  const int n = op.n, nb = op.nb, len = op.len;
  const float dt = op.dt, mass = op.mass, k_dm = op.k_dm;
  float2* v = D.fv;
  float2* X = D.cg_x;
  float2* R = D.cg_r;
  float2* Z = D.cg_z;
  float2* P = D.cg_p;
  float2* AP = D.cg_ap;
  float* INV = D.cg_inv;
  const uint32_t fc = S.force_cg;
  uint32_t visc_iters = 0;
  if (!op.joint) {
    double bn = sqrt(cg_dot(n, v, v));
    if (bn > 0.0 && isfinite(bn)) {
      double target = (double)1e-3f * bn;
      cg_apply(D, op, v, AP);
      for (int i = tid(); i < n; i += nthr()) { float2 r = vsub(v[i], AP[i]); R[i] = r; P[i] = r; }
      bsync();
      double rs_old = cg_dot(n, R, R);
      if (!(sqrt(rs_old) <= target)) {
        uint32_t maxit = fc ? fc : 50u;
        visc_iters = 50;
        for (uint32_t it = 1; it <= maxit; it++) {
          cg_apply(D, op, P, AP);
          double pap = cg_dot(n, P, AP);
          if (pap <= 0.0 || !isfinite(pap)) { visc_iters = it; break; }
          float alpha = (float)(rs_old / pap);
          for (int i = tid(); i < n; i += nthr()) {
            v[i] = vadd(v[i], smul(alpha, P[i]));
            R[i] = vsub(R[i], smul(alpha, AP[i]));
          }
          bsync();
          double rs_new = cg_dot(n, R, R);
          if (!fc && sqrt(rs_new) <= target) { visc_iters = it; break; }
          if (fc && it == maxit) { visc_iters = it; break; }
          float bta = (float)(rs_new / rs_old);
          for (int i = tid(); i < n; i += nthr()) P[i] = vadd(R[i], smul(bta, P[i]));
          bsync();
          rs_old = rs_new;
        }
      }
    }
  } else {
    // warm start x = [v + guess, V0, (W0, 0)]; Jacobi preconditioner; rhs b kept in P first
    const bool hg = sc->has_guess != 0u;
    for (int i = tid(); i < n; i += nthr()) {
      float2 vi = v[i];
      X[i] = hg ? vadd(vi, D.fguess[i]) : vi;
      D.lv[i] = vi;
      float sw = 0.0f;
      for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) sw += D.W[k];
      float dg = mass * (1.0f + dt * (0.5f * 4.0f * sw));
      if (op.ngroups)
        for (uint32_t q = D.gf_start[i]; q < D.gf_start[i + 1]; q++) dg += k_dm * D.g_c[D.gf_perm[q]];
      INV[i] = 1.0f / fmaxf(dg, 1e-20f);
      P[i] = smul(mass, vi);
    }
    for (int b = tid(); b < nb; b += nthr()) {
      float wb = D.bom[b] + D.ang[b] / D.binertia;
      D.bw[b] = wb;
      X[n + b] = D.bv0[b];
      X[n + nb + b] = v2(wb, 0.0f);
      float dv_ = D.bmass, dw_ = D.binertia;
      if (op.ngroups)
        for (uint32_t g = D.g_bstart[b]; g < D.g_bstart[b + 1]; g++) { dv_ += k_dm * D.g_c[g]; dw_ += k_dm * D.g_l2[g]; }
      if (op.nlinks)
        for (uint32_t q = D.li_start[b]; q < D.li_start[b + 1]; q++) {
          uint32_t l = D.rec_ord[D.li_perm[q]] >> 1;
          dv_ += dt * D.l_c[l];
        }
      INV[n + b] = 1.0f / fmaxf(dv_, 1e-20f);
      INV[n + nb + b] = 1.0f / fmaxf(dw_, 1e-20f);
      P[n + b] = smul(D.bmass, D.bv0[b]);
      P[n + nb + b] = v2(D.binertia * wb, 0.0f);
    }
    bsync();
    const double bn = cg_res_norm(len, INV, P);
    cg_apply(D, op, X, AP);
    for (int i = tid(); i < len; i += nthr()) {
      float2 r = vsub(P[i], AP[i]);
      R[i] = r;
      Z[i] = vmul(r, INV[i]);
    }
    bsync();
    uint32_t iters = 0;
    const double rn = cg_res_norm(len, INV, R);
    if (bn > 0.0 && isfinite(bn) && rn > (double)1e-3f * bn) {
      const double target = (double)1e-3f * bn;
      for (int i = tid(); i < len; i += nthr()) P[i] = Z[i];
      bsync();
      double rz_old = cg_dot(len, R, Z);
      uint32_t maxit = fc ? fc : 200u;
      for (uint32_t it = 1; it <= maxit; it++) {
        iters = it;
        cg_apply(D, op, P, AP);
        double pap = cg_dot(len, P, AP);
        if (pap <= 0.0 || !isfinite(pap)) break;
        float alpha = (float)(rz_old / pap);
        for (int i = tid(); i < len; i += nthr()) {
          X[i] = vadd(X[i], smul(alpha, P[i]));
          R[i] = vsub(R[i], smul(alpha, AP[i]));
        }
        bsync();
        if (!fc && cg_res_norm(len, INV, R) <= target) break;
        if (fc && it == maxit) break;
        for (int i = tid(); i < len; i += nthr()) Z[i] = vmul(R[i], INV[i]);
        bsync();
        double rz_new = cg_dot(len, R, Z);
        float bta = (float)(rz_new / rz_old);
        for (int i = tid(); i < len; i += nthr()) P[i] = vadd(Z[i], smul(bta, P[i]));
        bsync();
        rz_old = rz_new;
      }
    }
    visc_iters = iters;
    for (int i = tid(); i < n; i += nthr()) {
      float2 vn = X[i];
      v[i] = vn;
      float2 d = vsub(vn, D.lv[i]);
      D.fguess[i] = vfinite(d) ? d : v2(0.0f, 0.0f);
    }
    for (int b = tid(); b < nb; b += nthr()) {
      float2 dv = vsub(X[n + b], D.bv0[b]);
      float dw = X[n + nb + b].x - D.bw[b];
      if (vfinite(dv) && isfinite(dw)) {
        D.imp[b] = vadd(D.imp[b], smul(D.bmass, dv));
        D.ang[b] += D.binertia * dw;
      }
    }
    if (tid() == 0) sc->has_guess = 1u;
  }
  if (tid() == 0) sc->visc_iters = visc_iters;
}

// ---------------------------------------------------------------------------------------------
// k_fluid_tail: mean shear rate (pbf.rs:800-849), speed clamp (:987-996), impulse clamp
// (:2330-2345), viscous and clamp energy deltas.
// ---------------------------------------------------------------------------------------------
KERNEL k_fluid_tail(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!fluid_active(D, S)) return;
  MgScalars* sc = D.sc;
  const int n = D.nf, nb = D.nb;
  const float dt = S.dt, h = D.h, mass = D.fmass, br = D.br;
  float2* x = D.fx;
  float2* v = D.fv;
  double s = 0.0;
  uint32_t c = 0;
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = x[i], vi = v[i];
    float gxx = 0.0f, gxy = 0.0f, gyx = 0.0f, gyy = 0.0f;
    for (uint32_t k = D.noff[i]; k < D.noff[i + 1]; k++) {
      uint32_t j = D.nbr[k];
      float2 d = vsub(xi, x[j]);
      float r = vlen(d);
      float2 g = spiky_grad_k(d, r, h, D.k_spiky);
      float coeff = mass / fmaxf(D.dens[j], 1e-6f);
      float2 dv = vsub(v[j], vi);
      gxx += coeff * dv.x * g.x; gxy += coeff * dv.x * g.y;
      gyx += coeff * dv.y * g.x; gyy += coeff * dv.y * g.y;
    }
    float dxy = 0.5f * (gxy + gyx);
    float rate = sqrtf(2.0f * (gxx * gxx + gyy * gyy + 2.0f * dxy * dxy));
    if (isfinite(rate)) { s += (double)rate; c++; }
  }
  s = block_sum_d(s);
  c = block_sum_u(c);
  const double ke_after_visc = sum_ke(D, n);
  for (int i = tid(); i < n; i += nthr()) {
    float2 vi = v[i];
    float sp = vlen(vi);
    if (!isfinite(sp)) v[i] = v2(0.0f, 0.0f);
    else if (sp > D.v_max) v[i] = vmul(vi, D.v_max / sp);
  }
  bsync();
  const double ke_after_clamp = sum_ke(D, n);
  uint32_t hits = 0;
  if (nb > 0) {
    const float cl = 60.0f * D.bmass * 9.81f * dt;
    const float acl = cl * br;
    for (int b = tid(); b < nb; b += nthr()) {
      float2 im = D.imp[b];
      float m = vlen(im);
      if (m > cl && m > 1e-9f) { D.imp[b] = vmul(im, cl / m); hits++; }
      D.ang[b] = fclamp(D.ang[b], -acl, acl);
    }
  }
  hits = block_sum_u(hits);
  if (tid() == 0) {
    sc->mean_shear = c == 0 ? 0.0f : (float)(s / (double)c);
    sc->ke_visc = (float)(ke_after_visc - D.st->ke_mark);
    sc->ke_clamp = (float)(ke_after_clamp - ke_after_visc);
    sc->clamp_hits_total += hits;
  }
}

// ---------------------------------------------------------------------------------------------
// DEM (dem.rs:530-1066)
// ---------------------------------------------------------------------------------------------
// dem.rs:1069-1078. The wrap loops are bounded: a difference beyond 64 turns (or a non-finite
// one) only arises from an already diverged state, where the unbounded loop would never end.
DEV float angle_diff(float a, float b) {
  float d = a - b;
  for (int k = 0; k < 64 && d > 3.14159265f; k++) d -= 2.0f * 3.14159265f;
  for (int k = 0; k < 64 && d < -3.14159265f; k++) d += 2.0f * 3.14159265f;
  return d;
}
DEV float eff_friction(float mk, float ms, float vt, float vs) {  // dem.rs:128-138
  if (ms <= mk) return mk;
  float ratio = vt / vs;
  float w = expf(-(ratio * ratio));
  return mk + (ms - mk) * w;
}
DEV int impact_bin(float e) {  // dem.rs:99-109
  if (!isfinite(e) || e <= 0.0f) return -1;
  float ec = fclamp(e, 1e-6f, 1.0f);
  float lmin = log10f(1e-6f), lmax = 0.0f;
  float t = (log10f(ec) - lmin) / (lmax - lmin);
  uint32_t idx = sat_u32(t * 12.0f);
  return (int)min(idx, 11u);
}
DEV bool dem_active(const MgDrum& D, const MgSubstepIn& S) { return D.nb > 0 && S.dt > 0.0f; }

// k_dem_a: sanitize + predict with the coupling impulses (dem.rs:558-595), broadphase pairs in
// emission order (:626-633), swept wall CCD (:674-706).
KERNEL k_dem_a(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  MgScalars* sc = D.sc;
  MgStage* st = D.st;
  if (tid() == 0) {
    sc->dem_wall_work = 0.0f; sc->dem_dissipated = 0.0f; sc->dem_coupling_work = 0.0f; sc->collisions = 0;
    sc->n_pairs = 0; sc->n_impacts = 0;
    for (int k = 0; k < 12; k++) sc->dem_hist[k] = 0;
    st->npairs = 0;
  }
  if (!dem_active(D, S)) return;
  uint32_t* err = &sc->err;
  const int n = D.nb;
  const float dt = S.dt, r = D.br;
  float2* x = D.bx;
  float2* v = D.bv;
  float* th = D.bth;
  float* om = D.bom;
  const float w = D.bmass > 0.0f ? 1.0f / D.bmass : 0.0f;
  const float w_rot = D.binertia > 0.0f ? 1.0f / D.binertia : 0.0f;
  double cw = 0.0, e_pred = 0.0;
  float maxd = 0.0f, maxs = 0.0f;
  for (int i = tid(); i < n; i += nthr()) {
    float2 xi = x[i], vi = v[i];
    if (!vfinite(xi) || !vfinite(vi)) { xi = v2(0.0f, 0.0f); vi = v2(0.0f, 0.0f); }
    float oi = om[i], ti = th[i];
    D.bx0[i] = xi; D.bvpre[i] = vi; D.bth0[i] = ti; D.bompre[i] = oi;
    vi.y += GRAV * dt;
    double kb = 0.5 * (double)D.bmass * (double)vlen2(vi) + 0.5 * (double)D.binertia * (double)(oi * oi);
    vi = vadd(vi, vmul(D.imp[i], w));
    oi += D.ang[i] * w_rot;
    double ka = 0.5 * (double)D.bmass * (double)vlen2(vi) + 0.5 * (double)D.binertia * (double)(oi * oi);
    cw += ka - kb;
    xi = vadd(xi, vmul(vi, dt));
    ti += oi * dt;
    x[i] = xi; v[i] = vi; om[i] = oi; th[i] = ti;
    e_pred += (double)(0.5f * D.bmass * vlen2(vi));
    e_pred += (double)(0.5f * D.binertia * oi * oi);
    e_pred += (double)(D.bmass * 9.81f * xi.y);
    maxd = fmaxf(maxd, vlen(vi) * dt / fmaxf(2.0f * r, 1e-12f));
    maxs = fmaxf(maxs, vlen(vi));
  }
  cw = block_sum_d(cw);
  e_pred = block_sum_d(e_pred);
  maxd = block_max_f(maxd);
  maxs = block_max_f(maxs);
  const float margin = (float)max(D.dem_iters, 1u) * 0.2f * 2.0f * r;
  const float cell = fmaxf(fmaxf(2.0f * r * 1.05f, 2.0f * r + maxs * dt + margin), 1e-6f);
  GridView g = grid_build(x, n, cell, scratchC(D), err);
  uint32_t* pc = D.cursor;  // per-cell pair counts -> offsets
  const int nc = g.ncells;
  for (int k = tid(); k < nc; k += nthr()) {
    uint32_t s = g.start[k], e = g.start[k + 1];
    uint32_t cnt = 0;
    if (s != e) {
      int32_t cx, cy;
      if (g.dense) { cx = g.min_cx + (int32_t)((int64_t)k / g.H); cy = g.min_cy + (int32_t)((int64_t)k % g.H); }
      else { cx = key_cx(g.skeys[k]); cy = key_cy(g.skeys[k]); }
      uint32_t m = e - s;
      cnt = m * (m - 1) / 2;
      for (int o = 0; o < 4; o++) {
        const int ox = o == 3 ? -1 : (o == 2 ? 0 : 1);
        const int oy = o == 0 ? 0 : 1;
        int32_t ax = sadd_i32(cx, ox), ay = sadd_i32(cy, oy);
        if (ax == cx && ay == cy) continue;
        uint32_t os, oe;
        if (grid_range(g, ax, ay, os, oe)) cnt += m * (oe - os);
      }
    }
    pc[k] = cnt;
  }
  if (tid() == 0) pc[nc] = 0;
  bsync();
  block_scan_excl(pc, nc + 1);
  uint32_t npairs = nc ? pc[nc] : 0u;
  if (npairs > (uint32_t)D.pair_cap) { if (tid() == 0) atomicOr(err, MG_ERR_PAIRS); npairs = D.pair_cap; }
  for (int k = tid(); k < nc; k += nthr()) {
    uint32_t s = g.start[k], e = g.start[k + 1];
    if (s == e) continue;
    uint32_t wpos = pc[k];
    int32_t cx, cy;
    if (g.dense) { cx = g.min_cx + (int32_t)((int64_t)k / g.H); cy = g.min_cy + (int32_t)((int64_t)k % g.H); }
    else { cx = key_cx(g.skeys[k]); cy = key_cy(g.skeys[k]); }
    for (uint32_t a = s; a < e; a++)
      for (uint32_t b = a + 1; b < e; b++) {
        uint32_t i = g.idx[a], j = g.idx[b];
        if (wpos < (uint32_t)D.pair_cap) { D.pair_i[wpos] = min(i, j); D.pair_j[wpos] = max(i, j); }
        wpos++;
      }
    for (int o = 0; o < 4; o++) {
      const int ox = o == 3 ? -1 : (o == 2 ? 0 : 1);
      const int oy = o == 0 ? 0 : 1;
      int32_t ax = sadd_i32(cx, ox), ay = sadd_i32(cy, oy);
      if (ax == cx && ay == cy) continue;
      uint32_t os, oe;
      if (!grid_range(g, ax, ay, os, oe)) continue;
      for (uint32_t a = s; a < e; a++)
        for (uint32_t b = os; b < oe; b++) {
          uint32_t i = g.idx[a], j = g.idx[b];
          if (wpos < (uint32_t)D.pair_cap) { D.pair_i[wpos] = min(i, j); D.pair_j[wpos] = max(i, j); }
          wpos++;
        }
    }
  }
  for (uint32_t k = tid(); k < npairs; k += nthr()) D.lam_bb[k] = 0.0f;
  for (int i = tid(); i < n; i += nthr()) {
    float2 nrm;
    float ce = drum_sdf_world(D, x[i], S.cn, S.sn, S.cp, S.sp, nrm) - r;
    float t = 1.0f;
    D.lam_bw[i] = 0.0f;
    if (!(ce >= -0.2f * 2.0f * r)) {
      float tt = 1.0f;
      if (toi_swept(D, D.bx0[i], x[i], S.angle, S.angle_next, S.ca0, S.sa0, r, tt)) t = fminf(t, tt);
    }
    if (t < 1.0f) {
      float2 x0 = D.bx0[i];
      float pl = vlen(vsub(x[i], x0));
      float skin = 1e-2f * r;
      float ts = pl > 1e-9f ? skin / pl : 1.0f;
      float tc = fminf(t + ts, 1.0f);
      x[i] = vadd(x0, vmul(vsub(x[i], x0), tc));
    }
  }
  if (tid() == 0) {
    st->npairs = npairs;
    st->cw = cw;
    st->e_pred = e_pred;
    st->maxd = maxd;
  }
}

// k_dem_b: projection (dem.rs:708-756), friction (:762-890), restitution + impacts (:907-1019),
// rolling resistance (:1036-1055), energies and stats (:1057-1065).
KERNEL k_dem_b(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (!dem_active(D, S)) return;
  MgScalars* sc = D.sc;
  const MgStage* st = D.st;
  uint32_t* err = &sc->err;
  const int n = D.nb;
  const float dt = S.dt, r = D.br;
  float2* x = D.bx;
  float2* v = D.bv;
  float* th = D.bth;
  float* om = D.bom;
  const float w = D.bmass > 0.0f ? 1.0f / D.bmass : 0.0f;
  const float w_rot = D.binertia > 0.0f ? 1.0f / D.binertia : 0.0f;
  const uint32_t npairs = min(st->npairs, (uint32_t)D.pair_cap);
  // XPBD projection: Gauss-Seidel in pair emission order, then the walls.
  if (tid() == 0) {
    const float cmin = -0.2f * 2.0f * r;
    const float ws = w + w;
    for (uint32_t it = 0; it < max(D.dem_iters, 1u); it++) {
      for (uint32_t k = 0; k < npairs; k++) {
        uint32_t i = D.pair_i[k], j = D.pair_j[k];
        if (i >= (uint32_t)n || j >= (uint32_t)n) continue;
        float2 xi = x[i], xj = x[j];
        float2 d = vsub(xi, xj);
        float dist = vlen(d);
        float c = dist - 2.0f * r;
        if (c >= 0.0f) continue;
        float2 nh = dist > 1e-9f ? vdiv(d, dist) : v2(1.0f, 0.0f);
        if (ws <= 0.0f) continue;
        float ce = fmaxf(c, cmin);
        float dl = -ce / ws;
        x[i] = vadd(xi, smul(w * dl, nh));
        x[j] = vsub(xj, smul(w * dl, nh));
        D.lam_bb[k] += dl;
      }
      for (int i = 0; i < n; i++) {
        float2 nrm;
        float c = drum_sdf_world(D, x[i], S.cn, S.sn, S.cp, S.sp, nrm) - r;
        if (c >= 0.0f) continue;
        if (w <= 0.0f) continue;
        float ce = fmaxf(c, cmin);
        float dl = -ce / w;
        x[i] = vadd(x[i], smul(w * dl, nrm));
        D.lam_bw[i] += dl;
      }
    }
  }
  bsync();
  for (int i = tid(); i < n; i += nthr()) {
    v[i] = vdiv(vsub(x[i], D.bx0[i]), dt);
    om[i] = angle_diff(th[i], D.bth0[i]) / dt;
  }
  // contacts sorted by (i, j) (dem.rs:459-467); only positive-lambda contacts act below
  __shared__ uint32_t s_nct;
  if (tid() == 0) s_nct = 0;
  bsync();
  for (uint32_t k = tid(); k < npairs; k += nthr()) {
    if (D.lam_bb[k] > 0.0f) {
      uint32_t p = atomicAdd(&s_nct, 1u);
      if (p < (uint32_t)D.rec_cap) {
        D.rec_key[p] = D.pair_i[k];
        D.rec_ord[p] = D.pair_j[k];
        D.rec_v0[p] = __uint_as_float(k);
      }
    }
  }
  bsync();
  const uint32_t nct = min(s_nct, (uint32_t)D.rec_cap);
  ordered_csr((int)nct, D.rec_key, D.rec_ord, n, D.ct_start, D.ct_perm, D.cursor);
  const uint32_t nct_sorted = D.ct_start[n];
  if (tid() == 0) {
    const float wst = w + w + r * r * w_rot + r * r * w_rot;
    for (uint32_t q = 0; q < nct_sorted; q++) {
      uint32_t rr = D.ct_perm[q];
      uint32_t i = D.rec_key[rr], j = D.rec_ord[rr];
      uint32_t k = __float_as_uint(D.rec_v0[rr]);
      if (j >= (uint32_t)n || k >= npairs) continue;
      float lam = D.lam_bb[k];
      float2 d = vsub(x[i], x[j]);
      float dist = vlen(d);
      float2 nh = dist > 1e-9f ? vdiv(d, dist) : v2(1.0f, 0.0f);
      float2 t_ = v2(-nh.y, nh.x);
      float vt = vdot(vsub(v[i], v[j]), t_) - r * om[i] - r * om[j];
      if (wst <= 0.0f) continue;
      float raw = -vt * dt / wst;
      float mu = eff_friction(D.fric_bb, D.fric_bb_s, vt, D.fric_vscale);
      float mm = mu * lam;
      float dlt = fclamp(raw, -mm, mm);
      x[i] = vadd(x[i], smul(w * dlt, t_));
      x[j] = vsub(x[j], smul(w * dlt, t_));
      th[i] -= r * w_rot * dlt;
      th[j] -= r * w_rot * dlt;
      float2 dv = vdiv(smul(w * dlt, t_), dt);
      float dom = -r * w_rot * dlt / dt;
      v[i] = vadd(v[i], dv); v[j] = vsub(v[j], dv);
      om[i] += dom; om[j] += dom;
    }
  }
  bsync();
  double wwork = 0.0;
  {
    const float wst = w + r * r * w_rot;
    for (int i = tid(); i < n; i += nthr()) {
      float lam = D.lam_bw[i];
      if (lam <= 0.0f) continue;
      float2 nh;
      drum_sdf_world(D, x[i], S.cn, S.sn, S.cp, S.sp, nh);
      float2 t_ = v2(-nh.y, nh.x);
      float2 vw = wall_velocity(D.omega, vsub(x[i], smul(r, nh)));
      float vt = vdot(vsub(v[i], vw), t_) - r * om[i];
      if (wst <= 0.0f) continue;
      float raw = -vt * dt / wst;
      float mu = eff_friction(D.fric_bw, D.fric_bw_s, vt, D.fric_vscale);
      float mm = mu * lam;
      float dlt = fclamp(raw, -mm, mm);
      x[i] = vadd(x[i], smul(w * dlt, t_));
      th[i] -= r * w_rot * dlt;
      v[i] = vadd(v[i], vdiv(smul(w * dlt, t_), dt));
      om[i] += -r * w_rot * dlt / dt;
      wwork += (double)((dlt / dt) * vdot(t_, vw));
      wwork += (double)((lam / dt) * vdot(nh, vw));
    }
  }
  wwork = block_sum_d(wwork);
  for (int i = tid(); i < n; i += nthr()) {
    v[i] = vdiv(vsub(x[i], D.bx0[i]), dt);
    om[i] = angle_diff(th[i], D.bth0[i]) / dt;
  }
  bsync();
  __shared__ uint32_t s_coll, s_nimp;
  if (tid() == 0) {
    uint32_t coll = 0, nimp = 0;
    const float ws = w + w;
    for (uint32_t q = 0; q < nct_sorted; q++) {
      uint32_t rr = D.ct_perm[q];
      uint32_t i = D.rec_key[rr], j = D.rec_ord[rr];
      if (j >= (uint32_t)n) continue;
      float2 d = vsub(x[i], x[j]);
      float dist = vlen(d);
      float2 nh = dist > 1e-9f ? vdiv(d, dist) : v2(1.0f, 0.0f);
      float2 vpi = D.bvpre[i], vpj = D.bvpre[j];
      float vnp = vdot(vsub(vpi, vpj), nh);
      bool fresh = vnp < -0.02f;
      if (fresh) {
        coll++;
        int bin = impact_bin(0.5f * (D.bmass * 0.5f) * vnp * vnp);
        if (bin >= 0) sc->dem_hist[bin]++;
      }
      float e = fresh ? D.rest_bb : 0.0f;
      if (-vnp > D.impact_thr) {
        float2 t_ = v2(-nh.y, nh.x);
        float slip = vdot(vsub(vpi, vpj), t_) - r * D.bompre[i] - r * D.bompre[j];
        if (nimp < (uint32_t)D.imp_cap) {
          MgImpact rec; rec.kind = 0; rec.ball = i; rec.partner = j; rec.vn = -vnp; rec.slip = fabsf(slip); rec.e = e;
          D.impacts[nimp] = rec;
        }
        nimp++;
      }
      float target = fmaxf(fmaxf(-e * vnp, 0.0f), fmaxf(vnp, 0.0f));
      float vnn = vdot(vsub(v[i], v[j]), nh);
      float dvn = target - vnn;
      if (ws <= 0.0f) continue;
      float im = dvn / ws;
      v[i] = vadd(v[i], smul(w * im, nh));
      v[j] = vsub(v[j], smul(w * im, nh));
    }
    s_coll = coll;
    s_nimp = nimp;
  }
  bsync();
  // walls in ball order: each wall contact touches only its own ball; records keep ball order
  {
    __shared__ uint32_t s_wcnt[MG_BLOCK / 64];
    for (int base = 0; base < n; base += nthr()) {
      int i = base + tid();
      bool rec = false, fresh = false;
      MgImpact rc;
      rc.kind = 1; rc.ball = 0; rc.partner = MG_NO_SLOT; rc.vn = 0.0f; rc.slip = 0.0f; rc.e = 0.0f;
      if (i < n && D.lam_bw[i] > 0.0f) {
        float2 nh;
        drum_sdf_world(D, x[i], S.cn, S.sn, S.cp, S.sp, nh);
        float2 vw = wall_velocity(D.omega, vsub(x[i], smul(r, nh)));
        float2 vpi = D.bvpre[i];
        float vnp = vdot(vsub(vpi, vw), nh);
        fresh = vnp < -0.02f;
        if (fresh) {
          int bin = impact_bin(0.5f * D.bmass * vnp * vnp);
          if (bin >= 0) atomicAdd(&sc->dem_hist[bin], 1u);
        }
        float e = fresh ? D.rest_bw : 0.0f;
        if (-vnp > D.impact_thr) {
          float2 t_ = v2(-nh.y, nh.x);
          float slip = vdot(vsub(vpi, vw), t_) - r * D.bompre[i];
          rc.ball = (uint32_t)i; rc.vn = -vnp; rc.slip = fabsf(slip); rc.e = e;
          rec = true;
        }
        float target = fmaxf(fmaxf(-e * vnp, 0.0f), fmaxf(vnp, 0.0f));
        float vnn = vdot(vsub(v[i], vw), nh);
        float dvn = target - vnn;
        if (w > 0.0f) v[i] = vadd(v[i], smul(dvn, nh));
      }
      uint64_t m = __ballot(rec);
      uint64_t mf = __ballot(fresh);
      int lane = tid() & 63, wv = tid() >> 6;
      if (lane == 0) {
        s_wcnt[wv] = (uint32_t)__popcll(m);
        if (mf) atomicAdd(&s_coll, (uint32_t)__popcll(mf));
      }
      bsync();
      uint32_t off = s_nimp;
      for (int q = 0; q < wv; q++) off += s_wcnt[q];
      uint32_t pos = off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (rec && pos < (uint32_t)D.imp_cap) D.impacts[pos] = rc;
      bsync();
      if (tid() == 0) {
        uint32_t t = 0;
        for (int q = 0; q < nthr() / 64; q++) t += s_wcnt[q];
        s_nimp += t;
      }
      bsync();
    }
  }
  // rolling resistance
  if (D.rolling > 0.0f && D.binertia > 0.0f) {
    float* tl = D.toi;
    for (int i = tid(); i < n; i += nthr()) tl[i] = 0.0f;
    bsync();
    if (tid() == 0) {
      for (uint32_t q = 0; q < nct_sorted; q++) {
        uint32_t rr = D.ct_perm[q];
        uint32_t k = __float_as_uint(D.rec_v0[rr]);
        uint32_t i = D.rec_key[rr], j = D.rec_ord[rr];
        if (k >= npairs || j >= (uint32_t)n) continue;
        float lam = fmaxf(D.lam_bb[k], 0.0f);
        tl[i] += lam;
        tl[j] += lam;
      }
    }
    bsync();
    for (int i = tid(); i < n; i += nthr()) {
      float t = tl[i] + fmaxf(D.lam_bw[i], 0.0f);
      float o = om[i];
      if (t <= 0.0f || o == 0.0f) continue;
      float nf = t / dt;
      float mdo = D.rolling * nf * r / D.binertia * dt;
      float dom = fminf(mdo, fabsf(o));
      om[i] = o - (signbit(o) ? -1.0f : 1.0f) * dom;
    }
    bsync();
  }
  double e_fin = 0.0;
  for (int i = tid(); i < n; i += nthr()) {
    float2 vi = v[i];
    float oi = om[i];
    e_fin += (double)(0.5f * D.bmass * vlen2(vi));
    e_fin += (double)(0.5f * D.binertia * oi * oi);
    e_fin += (double)(D.bmass * 9.81f * x[i].y);
  }
  e_fin = block_sum_d(e_fin);
  if (tid() == 0) {
    sc->dem_wall_work = (float)wwork;
    sc->collisions = s_coll;
    sc->dem_dissipated = (float)(st->e_pred + wwork - e_fin);
    sc->max_disp = fmaxf(sc->max_disp, st->maxd);
    sc->dem_coupling_work = (float)st->cw;
    sc->n_pairs = npairs;
    sc->n_impacts = min(s_nimp, (uint32_t)D.imp_cap);
    if (s_nimp > (uint32_t)D.imp_cap) atomicOr(err, MG_ERR_IMPACTS);
  }
}

// ---------------------------------------------------------------------------------------------
// k_epilogue: energy budget (lib.rs:486-504), EMAs (:516-565), stress accumulator with the pair
// debounce window (stress.rs:110-171), drum clock (:507-508). Single lane, reference order.
// ---------------------------------------------------------------------------------------------
DEV uint64_t tbl_hash(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
  return k;
}
KERNEL k_epilogue(const MgDrum* __restrict__ drums, int slot) {
  const MgDrum& D = drums[blockIdx.x];
  const MgSubstepIn S = D.sched[slot];
  if (tid() != 0) return;
  MgScalars* sc = D.sc;
  const float dt = S.dt;
  const float fww = sc->fluid_wall_work, dww = sc->dem_wall_work;
  sc->eb_elapsed += (double)dt;
  sc->eb_dem_wall += (double)dww;
  sc->eb_fluid_wall += (double)fww;
  sc->eb_shaft += (double)(dww + fww);
  sc->eb_contact_diss += (double)sc->dem_dissipated;
  sc->eb_wall_slip += (double)(fww - sc->ke_noslip);
  sc->eb_viscous -= (double)sc->ke_visc;
  sc->eb_clamp_removed -= (double)sc->ke_clamp;
  sc->eb_interface += (double)(sc->dem_coupling_work + sc->ke_drag + sc->ke_buoy);
  sc->eb_fluid_wall_ang += (double)(float)sc->wall_ang;
  const float alpha = S.alpha;
  const float pinst = (dww + fww) / dt;
  sc->power_draw_w += alpha * (pinst - sc->power_draw_w);
  const float tinst = fabsf(D.omega) > 1e-6f ? pinst / D.omega : 0.0f;
  sc->torque_nm += alpha * (tinst - sc->torque_nm);
  const double window = D.window;
  const double now = S.now;
  sc->st_updates++;
  double c0 = 0, c1 = 0, c2 = 0, c3 = 0, e0 = 0, e1 = 0, e2 = 0, e3 = 0;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0;
  double* hist = D.st->hist;
  for (int k = 0; k < 28; k++) hist[k] = 0.0;
  const uint32_t cap = (uint32_t)D.tbl_cap;
  const uint32_t nimp = min(sc->n_impacts, (uint32_t)D.imp_cap);
  for (uint32_t q = 0; q < nimp; q++) {
    MgImpact rc = D.impacts[q];
    if (window > 0.0) {
      uint32_t a = rc.partner == MG_NO_SLOT ? rc.ball : min(rc.ball, rc.partner);
      uint32_t b = rc.partner == MG_NO_SLOT ? MG_NO_SLOT : max(rc.ball, rc.partner);
      uint64_t key = ((uint64_t)a << 32) | b;
      uint32_t hh = (uint32_t)(tbl_hash(key) & (uint64_t)(cap - 1));
      int32_t found = -1, first_free = -1;
      for (uint32_t p = 0; p < cap; p++) {
        uint32_t s = (hh + p) & (cap - 1);
        uint64_t k = D.tbl_key[s];
        if (k == ~0ull) { if (first_free < 0) first_free = (int32_t)s; break; }
        if (k == key) { found = (int32_t)s; break; }
        if (first_free < 0 && !(now - D.tbl_time[s] < window)) first_free = (int32_t)s;
      }
      if (found >= 0) {
        if (now - D.tbl_time[found] < window) continue;
        D.tbl_time[found] = now;
      } else if (first_free >= 0) {
        if (D.tbl_key[first_free] == ~0ull) sc->st_table_used++;
        D.tbl_key[first_free] = key;
        D.tbl_time[first_free] = now;
      } else {
        atomicOr(&sc->err, MG_ERR_TABLE);
      }
    }
    const bool bb = rc.kind == 0;
    const double vn = rc.vn, vs = rc.slip;
    const int cls = bb ? (vs > vn ? 1 : 0) : (vs > vn ? 3 : 2);
    const double m_red = bb ? 0.5 * D.m3 : D.m3;
    const double se = 0.5 * m_red * (vn * vn + vs * vs);
    const double eh = se / D.scale_j;
    if (cls == 0) { c0 += 1.0; e0 += eh; } else if (cls == 1) { c1 += 1.0; e1 += eh; }
    else if (cls == 2) { c2 += 1.0; e2 += eh; } else { c3 += 1.0; e3 += eh; }
    int bin = 0;
    if (!(isnan(eh) || eh <= 0.0)) {
      double t = (log10(eh) - -3.0) * 4.0;
      double tm = fmax(t, 0.0);
      bin = tm >= 27.0 ? 27 : (int)tm;
    }
    hist[bin] += 1.0;
    if (eh >= 0.03) a0 += se;
    if (eh >= 0.1) a1 += se;
    if (eh >= 0.3) a2 += se;
    if (eh >= 1.0) a3 += se;
    if (eh >= 3.0) a4 += se;
    if (eh >= 10.0) a5 += se;
  }
  const double wall_work = (double)(dww + fww);
  sc->st_cum_wall += wall_work;
  sc->st_cum_above[0] += a0; sc->st_cum_above[1] += a1; sc->st_cum_above[2] += a2;
  sc->st_cum_above[3] += a3; sc->st_cum_above[4] += a4; sc->st_cum_above[5] += a5;
  const double aa = (double)alpha, sdt = (double)dt;
  sc->st_count[0] += aa * (c0 / sdt - sc->st_count[0]);
  sc->st_se[0] += aa * (e0 / sdt - sc->st_se[0]);
  sc->st_count[1] += aa * (c1 / sdt - sc->st_count[1]);
  sc->st_se[1] += aa * (e1 / sdt - sc->st_se[1]);
  sc->st_count[2] += aa * (c2 / sdt - sc->st_count[2]);
  sc->st_se[2] += aa * (e2 / sdt - sc->st_se[2]);
  sc->st_count[3] += aa * (c3 / sdt - sc->st_count[3]);
  sc->st_se[3] += aa * (e3 / sdt - sc->st_se[3]);
  for (int k = 0; k < 28; k++) sc->st_hist[k] += aa * (hist[k] / sdt - sc->st_hist[k]);
  sc->st_above[0] += aa * (a0 / sdt - sc->st_above[0]);
  sc->st_above[1] += aa * (a1 / sdt - sc->st_above[1]);
  sc->st_above[2] += aa * (a2 / sdt - sc->st_above[2]);
  sc->st_above[3] += aa * (a3 / sdt - sc->st_above[3]);
  sc->st_above[4] += aa * (a4 / sdt - sc->st_above[4]);
  sc->st_above[5] += aa * (a5 / sdt - sc->st_above[5]);
  sc->collision_rate_per_s = (float)(sc->st_count[0] + sc->st_count[1] + sc->st_count[2] + sc->st_count[3]);
  const float dpi = fmaxf(sc->dem_dissipated, 0.0f) / dt;
  sc->dissipated_power_w += alpha * (dpi - sc->dissipated_power_w);
  // table hygiene: once half the slots are in use, keep only entries inside the debounce window
  if (window > 0.0 && sc->st_table_used > cap / 2) {
    uint32_t live = 0;
    for (uint32_t s = 0; s < cap; s++) {
      uint64_t k = D.tbl_key[s];
      if (k == ~0ull) continue;
      if (now - D.tbl_time[s] < window) { D.tbl_key[cap + live] = k; D.tbl_time[cap + live] = D.tbl_time[s]; live++; }
      D.tbl_key[s] = ~0ull;
    }
    for (uint32_t q = 0; q < live; q++) {
      uint64_t key = D.tbl_key[cap + q];
      uint32_t hh = (uint32_t)(tbl_hash(key) & (uint64_t)(cap - 1));
      for (uint32_t p = 0; p < cap; p++) {
        uint32_t s = (hh + p) & (cap - 1);
        if (D.tbl_key[s] == ~0ull) { D.tbl_key[s] = key; D.tbl_time[s] = D.tbl_time[cap + q]; break; }
      }
    }
    sc->st_table_used = live;
  }
  float ang = S.angle + D.omega * dt;
  float rmd = fmodf(ang, 6.28318531f);
  sc->drum_angle = rmd < 0.0f ? rmd + 6.28318531f : rmd;
  sc->sim_time = now + (double)dt;
}

// Launches the stage chain of one substep for every drum of the batch.
extern "C" hipError_t mg_launch_substep(const MgDrum* d_drums, int n_drums, int slot, hipStream_t stream) {
  const dim3 grid(n_drums), block(MG_BLOCK);
  hipLaunchKernelGGL(k_fluid_prep, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_density, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_couple, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_wall, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_drag, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_buoy, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_links, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_visc, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_fluid_tail, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_dem_a, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_dem_b, grid, block, 0, stream, d_drums, slot);
  hipLaunchKernelGGL(k_epilogue, grid, block, 0, stream, d_drums, slot);
  return hipGetLastError();
}
