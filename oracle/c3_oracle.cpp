// CPU oracle for the kernel-isolation workload: a periodic box of PBF particles.
//
// TEST INFRASTRUCTURE ONLY (included by mg_oracle.cpp; same rules as that file).
//
// The reference has no periodic domain, so this workload has no reference trace of its own. It
// applies the reference's fluid arithmetic, unchanged, to a box of side L with minimum-image pair
// vectors and wrapped positions:
//   predict                 x0 = x; x += v dt (gravity 0)             pbf.rs:1225-1235
//   neighbour lists         |d|^2 <= h^2, mutual, reference order      pbf.rs:262-288 (order as the
//                           grid.rs:291-362 emission gives it per particle, see c3_neighbors)
//   density passes          rho, one-sided C, lambda, dp, Jacobi       pbf.rs:1257-1334
//   velocity                v = (x - x0) / dt                          pbf.rs:1656-1674
//   density recompute       at the final positions                     pbf.rs:2136-2144
//   Morris weights          c_ij, unit directions                      pbf.rs:312-344
//   implicit viscosity CG   (I + dt L) v = v, warm start v             pbf.rs:359-480
// with exactly `passes` density passes, or (adaptive) the reference's count: at least
// max(passes, 1), stopping once max C <= 0.02, at most 6 more (pbf.rs:1261-1333); the CG stopping
// at `tol`·|b| after at most `max_cg` iterations (the reference uses 1e-3 and 50). Options the
// drum step has and the box leaves off by default: gravity in the predict (pbf.rs:1225-1235), the
// step-0 reset of non-finite particles (pbf.rs:1200-1205, always on), the fluid speed clamp after
// the CG (pbf.rs:987-996, :2321) and f64 kinetic-energy snapshots around the CG and the clamp
// (pbf.rs:924-926).
//
// Workload definition (synthetic; documented beside the GPU path):
//   side x side square lattice, spacing dx, kernel radius h = 2 dx, L = side·dx;
//   particle i = iy·side + ix at ((ix + 0.5) dx + jx, (iy + 0.5) dx + jy), jitter U(-0.1dx, 0.1dx);
//   velocity per component 0.01·N(0, 1) by Box-Muller (cosine branch), computed in f64;
//   one Rng(seed) stream in particle order: jx, jy, u1, u2, u3, u4;
//   mass = 1.02·rest / Σ_lattice W (the unjittered lattice is 2 % over-dense; density_ratio
//   replaces the 1.02).

namespace orc {

struct C3Config {
  uint32_t side = 1024;
  float dx = 1.0e-3f;
  float rest = 1000.0f;
  float mu = 1.0e-3f;
  float dt = 1.0f / 1440.0f;
  uint32_t passes = 4;
  float tol = 1.0e-6f;
  uint32_t max_cg = 500;
  uint64_t seed = 1;
  float yield_stress = 0.0f;  // Bingham yield stress (Pa); 0: Newtonian
  uint32_t adaptive = 0;      // 1: the reference's adaptive pass count with base max(passes, 1)
  float gravity = 0.0f;       // added to v.y * dt in the predict (the reference's GRAVITY is -9.81)
  float speed_clamp = 0.0f;   // fluid speed clamp v_max after the CG (pbf.rs:987-996); 0: off
  float density_ratio = 0.0f; // density of the unjittered lattice over rest (0: 1.02)
};

// Box geometry shared by the seeding and the step (all f32, as the GPU path computes it).
struct C3Box {
  float dx, h, L, halfL, cw, mass;
  uint32_t ncx;
  static C3Box make(const C3Config& c) {
    C3Box b;
    b.dx = c.dx;
    b.h = 2.0f * c.dx;
    b.L = (float)c.side * c.dx;
    b.halfL = 0.5f * b.L;
    // cells of width cw = L / ncx >= h, so every pair within h lies in adjacent cells
    uint32_t n = (uint32_t)std::floor((double)b.L / (double)b.h);
    while (n > 1 && b.L / (float)n < b.h) n--;
    b.ncx = std::max(n, 1u);
    b.cw = b.L / (float)b.ncx;
    float lw = 0.0f;  // Σ over the unjittered lattice around one site, row-major offsets
    for (int q = -2; q <= 2; q++)
      for (int p = -2; p <= 2; p++) {
        float rx = (float)p * c.dx, ry = (float)q * c.dx;
        lw += poly6<float>(rx * rx + ry * ry, b.h);
      }
    const float ratio = c.density_ratio > 0.0f ? c.density_ratio : 1.02f;
    b.mass = ratio * c.rest / lw;
    return b;
  }
  // rem_euclid (f32), as the reference wraps angles
  float wrap(float a) const { return rem_euclid<float>(a, L); }
  double wrap(double a) const { return rem_euclid<double>(a, (double)L); }
  template <class R> R mi(R d) const {  // minimum image of one component
    if (d > (R)halfL) return d - (R)L;
    if (d < -(R)halfL) return d + (R)L;
    return d;
  }
  template <class R> uint32_t cell1(R a) const {
    R q = std::floor(a / (R)cw);
    int64_t c = q < (R)0 ? 0 : (int64_t)q;
    return (uint32_t)std::min<int64_t>(c, (int64_t)ncx - 1);
  }
};

// Initial state in f32 (every precision policy starts from the same f32 values).
static void c3_seed(const C3Config& c, const C3Box& b, std::vector<float>& x, std::vector<float>& v) {
  const size_t n = (size_t)c.side * c.side;
  x.assign(2 * n, 0.0f);
  v.assign(2 * n, 0.0f);
  Rng rng(c.seed);
  const float jit = 0.1f * c.dx;
  const double two_pi = 6.283185307179586476925;
  for (uint32_t iy = 0; iy < c.side; iy++)
    for (uint32_t ix = 0; ix < c.side; ix++) {
      size_t i = (size_t)iy * c.side + ix;
      float jx = (rng.next_f32() * 2.0f - 1.0f) * jit;
      float jy = (rng.next_f32() * 2.0f - 1.0f) * jit;
      double u1 = rng.next_f32(), u2 = rng.next_f32(), u3 = rng.next_f32(), u4 = rng.next_f32();
      double g1 = std::sqrt(-2.0 * std::log(std::max(u1, 1e-12))) * std::cos(two_pi * u2);
      double g2 = std::sqrt(-2.0 * std::log(std::max(u3, 1e-12))) * std::cos(two_pi * u4);
      x[2 * i] = b.wrap(((float)ix + 0.5f) * c.dx + jx);
      x[2 * i + 1] = b.wrap(((float)iy + 0.5f) * c.dx + jy);
      v[2 * i] = (float)(0.01 * g1);
      v[2 * i + 1] = (float)(0.01 * g2);
    }
}

// Per-particle neighbour order of the reference (pbf.rs:262-288 over grid.rs:291-362): a point in
// cell c receives its pairs from the cells (-1,-1), (-1,0), (0,-1), then c itself (ascending
// index, skipping the point), then (1,0), (1,1), (0,1), (-1,1) and last (1,-1). With periodic
// cells the offsets wrap; ncx >= 3 keeps the nine cells distinct.
static const int kC3Ox[9] = {-1, -1, 0, 0, 1, 1, 0, -1, 1};
static const int kC3Oy[9] = {-1, 0, -1, 0, 0, 1, 1, 1, -1};

template <class R>
static Nbrs c3_neighbors(const std::vector<V2<R>>& x, const C3Box& b) {
  const size_t n = x.size();
  const uint32_t nc = b.ncx;
  // counting sort by cell, ascending index inside a cell
  std::vector<uint32_t> start((size_t)nc * nc + 1, 0), idx(n);
  std::vector<uint32_t> key(n);
  for (size_t i = 0; i < n; i++) {
    key[i] = b.cell1(x[i].y) * nc + b.cell1(x[i].x);
    start[key[i] + 1]++;
  }
  for (size_t k = 1; k < start.size(); k++) start[k] += start[k - 1];
  std::vector<uint32_t> cur(start.begin(), start.end() - 1);
  for (size_t i = 0; i < n; i++) idx[cur[key[i]]++] = (uint32_t)i;
  Nbrs L;
  L.off.assign(n + 1, 0);
  const R h2 = (R)b.h * (R)b.h;
  std::vector<uint32_t> tmp;
  for (size_t i = 0; i < n; i++) {
    int cx = (int)b.cell1(x[i].x), cy = (int)b.cell1(x[i].y);
    for (int o = 0; o < 9; o++) {
      uint32_t ax = (uint32_t)((cx + kC3Ox[o] + (int)nc) % (int)nc);
      uint32_t ay = (uint32_t)((cy + kC3Oy[o] + (int)nc) % (int)nc);
      uint32_t k = ay * nc + ax;
      for (uint32_t a = start[k]; a < start[k + 1]; a++) {
        uint32_t j = idx[a];
        if (j == (uint32_t)i) continue;
        V2<R> d(b.mi(x[i].x - x[j].x), b.mi(x[i].y - x[j].y));
        if (d.len2() <= h2) L.nb.push_back(j);
      }
    }
    L.off[i + 1] = (uint32_t)L.nb.size();
  }
  return L;
}

// Per-particle shear rate sqrt(2 D:D) from the SPH velocity gradient over the neighbour lists
// (pbf.rs:821-849), with the minimum-image pair vector of the periodic box.
template <class R>
static std::vector<R> c3_shear_rates(const std::vector<V2<R>>& x, const std::vector<V2<R>>& v, const Nbrs& nbr,
                                     const std::vector<R>& density, R mass, R h, const C3Box& b) {
  const size_t n = x.size();
  std::vector<R> out(n, (R)0);
  for (size_t i = 0; i < n; i++) {
    R gxx = 0, gxy = 0, gyx = 0, gyy = 0;
    for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
      const uint32_t j = nbr.nb[k];
      const V2<R> d(b.mi(x[i].x - x[j].x), b.mi(x[i].y - x[j].y));
      const R r = d.len();
      const V2<R> g = spiky_grad(d, r, h);
      const R coeff = mass / rmax(density[j], F(1e-6));
      const V2<R> dv = v[j] - v[i];
      gxx += coeff * dv.x * g.x;
      gxy += coeff * dv.x * g.y;
      gyx += coeff * dv.y * g.x;
      gyy += coeff * dv.y * g.y;
    }
    const R dxy = F(0.5) * (gxy + gyx);
    out[i] = std::sqrt(F(2.0) * (gxx * gxx + gyy * gyy + F(2.0) * dxy * dxy));
  }
  return out;
}
// Papanastasiou-regularised Bingham viscosity mu + tau_y (1 - exp(-g/g0)) / g, g0 = tau_y / (100 mu),
// with its g -> 0 limit below 1e-4 g0 and non-finite rates taken as 0 (pbf.rs:854-868, :490).
template <class R> static R c3_bingham(R g, R mu, R tau_y) {
  const R g0 = tau_y / (F(100.0) * mu);
  if (!std::isfinite(g)) g = (R)0;
  const R extra = g < F(1e-4) * g0 ? tau_y / g0 : tau_y * (F(1.0) - std::exp(-g / g0)) / g;
  return mu + extra;
}

// One density pass over the lists (pbf.rs:1257-1334): density, one-sided C, lambda and the
// Jacobi displacement dp at positions x (not moved here). gcache holds each edge's spiky gradient,
// as the reference's per-edge cache does (pbf.rs:1268-1285).
template <class R>
static R c3_density_pass(const std::vector<V2<R>>& x, const Nbrs& nbr, const C3Box& b, R mass, R rest,
                         std::vector<R>& density, std::vector<R>& lambda, std::vector<V2<R>>& dp,
                         std::vector<V2<R>>& gcache) {
  R max_c = 0;  // the largest constraint of the pass (f32::max from 0, NaN-ignoring)
  const size_t n = x.size();
  const R h = (R)b.h;
  density.assign(n, (R)0);
  lambda.assign(n, (R)0);
  dp.assign(n, V2<R>());
  gcache.resize(nbr.nb.size());
  for (size_t i = 0; i < n; i++) {
    R rho = mass * poly6((R)0, h);
    V2<R> gs;
    R sgs = 0;
    for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
      uint32_t j = nbr.nb[k];
      V2<R> d(b.mi(x[i].x - x[j].x), b.mi(x[i].y - x[j].y));
      R r2 = d.len2();
      rho += mass * poly6(r2, h);
      R r = d.len();
      V2<R> rg = spiky_grad(d, r, h);
      gcache[k] = rg;
      V2<R> g = rg * (mass / rest);
      gs += g;
      sgs += g.len2();
    }
    density[i] = rho;
    R c = rmax(density[i] / rest - F(1.0), (R)0);
    max_c = rmax(max_c, c);
    sgs += gs.len2();
    lambda[i] = -c / (sgs + F(200.0));
  }
  for (size_t i = 0; i < n; i++) {
    V2<R> s;
    for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) s += (lambda[i] + lambda[nbr.nb[k]]) * gcache[k];
    dp[i] = s * (mass / rest);
  }
  return max_c;
}

// kinetic_energy_f64 (pbf.rs:924-926): 0.5 m times the f64 sum of the f32 |v|^2.
template <class R> static double c3_kinetic_energy(const std::vector<V2<R>>& v, R mass) {
  double s = 0.0;
  for (const V2<R>& a : v) s += (double)a.len2();
  return 0.5 * (double)mass * s;
}
// clamp_fluid_speed (pbf.rs:987-996).
template <class R> static void c3_clamp_speed(std::vector<V2<R>>& v, R v_max) {
  for (V2<R>& a : v) {
    const R speed = a.len();
    if (!std::isfinite(speed)) a = V2<R>();
    else if (speed > v_max) a = a * (v_max / speed);
  }
}

// Density at positions x over the lists (pbf.rs:2136-2144).
template <class R>
static std::vector<R> c3_density_at(const std::vector<V2<R>>& x, const Nbrs& nbr, const C3Box& b, R mass) {
  const size_t n = x.size();
  const R h = (R)b.h;
  std::vector<R> density(n);
  for (size_t i = 0; i < n; i++) {
    R rho = mass * poly6((R)0, h);
    for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
      const uint32_t j = nbr.nb[k];
      rho += mass * poly6(V2<R>(b.mi(x[i].x - x[j].x), b.mi(x[i].y - x[j].y)).len2(), h);
    }
    density[i] = rho;
  }
  return density;
}

// Morris weights c_ij and unit directions of every edge (pbf.rs:312-344).
template <class R>
static void c3_morris(const std::vector<V2<R>>& x, const Nbrs& nbr, const std::vector<R>& density,
                      const std::vector<R>& mu_eff, const C3Box& b, R mass, std::vector<R>& W,
                      std::vector<V2<R>>& D) {
  const size_t n = x.size();
  const R h = (R)b.h;
  const R eta2 = F(0.01) * h * h;
  W.assign(nbr.nb.size(), (R)0);
  D.assign(nbr.nb.size(), V2<R>());
  for (size_t i = 0; i < n; i++) {
    R rho_i = rmax(density[i], F(1e-6));
    for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
      uint32_t j = nbr.nb[k];
      R rho_j = rmax(density[j], F(1e-6));
      V2<R> d(b.mi(x[i].x - x[j].x), b.mi(x[i].y - x[j].y));
      R r = d.len();
      V2<R> g = spiky_grad(d, r, h);
      R num = (mass * (mu_eff[i] + mu_eff[j]) / (rho_i * rho_j)) * std::fabs(d.dot(g));
      W[k] = num / (d.len2() + eta2);
      D[k] = r > F(1e-9) ? d / r : V2<R>(1, 0);
    }
  }
}

// Central-force viscous operator (L v)_i = sum_j 4 c_ij e_ij (e_ij . (v_i - v_j)) (pbf.rs:359-377,
// CENTRAL_VISCOSITY_FACTOR = 4 at :384).
template <class R>
static void c3_laplacian(const Nbrs& nbr, const std::vector<R>& W, const std::vector<V2<R>>& D,
                         const std::vector<V2<R>>& vv, std::vector<V2<R>>& out) {
  const size_t n = vv.size();
  out.resize(n);
  for (size_t i = 0; i < n; i++) {
    V2<R> s;
    for (uint32_t k = nbr.off[i]; k < nbr.off[i + 1]; k++) {
      V2<R> e = D[k];
      s += (F(4.0) * W[k] * e.dot(vv[i] - vv[nbr.nb[k]])) * e;
    }
    out[i] = s;
  }
}

template <class R, class A> static A c3_dotv(const std::vector<V2<R>>& a, const std::vector<V2<R>>& b) {  // pbf.rs:292-294
  A s = 0;
  for (size_t i = 0; i < a.size(); i++) s += (A)a[i].x * (A)b[i].x + (A)a[i].y * (A)b[i].y;
  return s;
}

// solve_implicit_viscosity (pbf.rs:423-480): CG on (I + dt L) v = v with warm start v, stopping
// at tol * |b| after at most max_cg iterations; dots and norms in A. Returns the iterations run;
// res and bnorm receive the last residual norm and |b|.
template <class R, class A>
static uint32_t c3_cg(std::vector<V2<R>>& v, const Nbrs& nbr, const std::vector<R>& W, const std::vector<V2<R>>& D,
                      R dt, float tol, uint32_t max_cg, A& res, A& bnorm) {
  const size_t n = v.size();
  A bn = std::sqrt(c3_dotv<R, A>(v, v));
  bnorm = bn;
  res = 0;
  if (bn <= (A)0 || !std::isfinite(bn)) return 0;
  A target = (A)tol * bn;
  std::vector<V2<R>> lv(n), r(n), p(n), ap(n);
  c3_laplacian(nbr, W, D, v, lv);
  for (size_t i = 0; i < n; i++) r[i] = v[i] - (v[i] + lv[i] * dt);
  p = r;
  A rs_old = c3_dotv<R, A>(r, r);
  res = std::sqrt(rs_old);
  if (std::sqrt(rs_old) <= target) return 0;
  uint32_t iters = 0;
  for (uint32_t it = 1; it <= max_cg; it++) {
    iters = it;
    c3_laplacian(nbr, W, D, p, lv);
    for (size_t i = 0; i < n; i++) ap[i] = p[i] + lv[i] * dt;
    A pap = c3_dotv<R, A>(p, ap);
    if (pap <= (A)0 || !std::isfinite(pap)) return iters;
    R alpha = (R)(rs_old / pap);
    for (size_t i = 0; i < n; i++) { v[i] += p[i] * alpha; r[i] -= ap[i] * alpha; }
    A rs_new = c3_dotv<R, A>(r, r);
    res = std::sqrt(rs_new);
    if (std::sqrt(rs_new) <= target) return iters;
    R bta = (R)(rs_new / rs_old);
    for (size_t i = 0; i < n; i++) p[i] = r[i] + p[i] * bta;
    rs_old = rs_new;
  }
  return iters;
}

template <class R, class A> struct C3Sim {
  C3Config cfg;
  C3Box box;
  std::vector<V2<R>> x, v;
  Nbrs nbr;                 // lists of the last step (at the predicted positions)
  std::vector<R> dens;      // density at the final positions of the last step
  uint32_t last_cg = 0;
  A last_res = 0, last_bnorm = 0;
  uint32_t last_passes = 0;                           // density passes of the last step
  double ke_before_cg = 0, ke_after_cg = 0, ke_after_clamp = 0;  // f64 snapshots of the last step

  void init(const C3Config& c) {
    cfg = c;
    box = C3Box::make(c);
    std::vector<float> xf, vf;
    c3_seed(c, box, xf, vf);
    size_t n = xf.size() / 2;
    x.resize(n);
    v.resize(n);
    for (size_t i = 0; i < n; i++) {
      x[i] = V2<R>((R)xf[2 * i], (R)xf[2 * i + 1]);
      v[i] = V2<R>((R)vf[2 * i], (R)vf[2 * i + 1]);
    }
  }
  V2<R> mi(V2<R> d) const { return V2<R>(box.mi(d.x), box.mi(d.y)); }
  V2<R> wrap(V2<R> p) const { return V2<R>(box.wrap(p.x), box.wrap(p.y)); }

  // Mean shear rate of the last step (pbf.rs:800-818 at :2303-2304: final positions, the step's
  // velocities, lists and recomputed densities); the sum in the accumulation type A.
  double mean_shear() const {
    if (nbr.off.empty()) return 0.0;
    const std::vector<R> rates = c3_shear_rates(x, v, nbr, dens, (R)box.mass, (R)box.h, box);
    A sum = 0;
    uint32_t c = 0;
    for (R g : rates)
      if (std::isfinite(g)) { sum += (A)g; c++; }
    return c == 0 ? 0.0 : (double)(R)(sum / (A)c);
  }

  void step() {
    const size_t n = x.size();
    const R dt = (R)cfg.dt, rest = (R)cfg.rest, mass = (R)box.mass, mu = (R)cfg.mu;
    // step 0: a non-finite particle restarts at rest at the origin (pbf.rs:1200-1205)
    for (size_t i = 0; i < n; i++)
      if (!x[i].finite() || !v[i].finite()) { x[i] = V2<R>(); v[i] = V2<R>(); }
    // predict (pbf.rs:1225-1235), positions wrapped into the box
    std::vector<V2<R>> x0 = x;
    const R gdt = (R)cfg.gravity * dt;
    for (size_t i = 0; i < n; i++) {
      v[i].y += gdt;
      x[i] = wrap(x[i] + v[i] * dt);
    }
    // neighbour lists at the predicted positions (pbf.rs:1238-1239)
    nbr = c3_neighbors(x, box);
    // density passes (pbf.rs:1257-1334): exactly cfg.passes, or the reference's adaptive count
    std::vector<R> density, lambda;
    std::vector<V2<R>> dp, gcache;
    const uint32_t base = cfg.adaptive ? std::max(cfg.passes, 1u) : cfg.passes;
    uint32_t count = 0;
    while (count < (cfg.adaptive ? base + 6u : base)) {
      const R max_c = c3_density_pass(x, nbr, box, mass, rest, density, lambda, dp, gcache);
      for (size_t i = 0; i < n; i++) x[i] = wrap(x[i] + dp[i]);
      count++;
      if (cfg.adaptive && count >= base && max_c <= F(0.02)) break;
    }
    last_passes = count;
    // velocity (pbf.rs:1656-1674) with the minimum-image displacement
    for (size_t i = 0; i < n; i++) v[i] = mi(x[i] - x0[i]) / dt;
    // density recompute (pbf.rs:2136-2144)
    density = c3_density_at(x, nbr, box, mass);
    dens = density;
    last_cg = 0;
    ke_before_cg = c3_kinetic_energy(v, mass);
    viscosity(mu);
    ke_after_cg = c3_kinetic_energy(v, mass);
    if (cfg.speed_clamp > 0.0f) c3_clamp_speed(v, (R)cfg.speed_clamp);
    ke_after_clamp = c3_kinetic_energy(v, mass);
  }

  // Implicit viscosity of the step's final state (pbf.rs:2136-2299, fluid-only form).
  void viscosity(R mu) {
    const size_t n = x.size();
    const R dt = (R)cfg.dt, h = (R)box.h, mass = (R)box.mass;
    const std::vector<R>& density = dens;
    if (!(mu > (R)0)) return;
    // effective viscosity per particle: Bingham from the shear rates at the recomputed densities
    // and the updated velocities (pbf.rs:2147-2151), else mu everywhere
    std::vector<R> mu_eff(n, mu);
    if (cfg.yield_stress > 0.0f) {
      const std::vector<R> rates = c3_shear_rates(x, v, nbr, density, mass, h, box);
      for (size_t i = 0; i < n; i++) mu_eff[i] = c3_bingham(rates[i], mu, (R)cfg.yield_stress);
    }
    std::vector<R> W;
    std::vector<V2<R>> D;
    c3_morris(x, nbr, density, mu_eff, box, mass, W, D);
    last_cg = c3_cg<R, A>(v, nbr, W, D, dt, cfg.tol, cfg.max_cg, last_res, last_bnorm);
  }
};

}  // namespace orc

// ---- C API (ctypes) for the kernel-isolation oracle. mode: 1 = mixed (f32, f64 dots), 2 = fp64.
struct C3Handle {
  int mode;
  void* p;
};
static orc::C3Config c3_cfg_from(const double* c) {
  orc::C3Config k;
  k.side = (uint32_t)c[0]; k.dx = (float)c[1]; k.rest = (float)c[2]; k.mu = (float)c[3]; k.dt = (float)c[4];
  k.passes = (uint32_t)c[5]; k.tol = (float)c[6]; k.max_cg = (uint32_t)c[7]; k.seed = (uint64_t)c[8];
  k.yield_stress = (float)c[9];
  k.adaptive = (uint32_t)c[10]; k.gravity = (float)c[11]; k.speed_clamp = (float)c[12];
  k.density_ratio = (float)c[13];
  return k;
}
#define C3DISPATCH(h, ...)                                                                   \
  do {                                                                                       \
    if ((h)->mode == 2) { auto* s = (orc::C3Sim<double, double>*)(h)->p; __VA_ARGS__; }      \
    else if ((h)->mode == 0) { auto* s = (orc::C3Sim<float, float>*)(h)->p; __VA_ARGS__; }   \
    else { auto* s = (orc::C3Sim<float, double>*)(h)->p; __VA_ARGS__; }                      \
  } while (0)

extern "C" {
// cfg: side, dx, rest, mu, dt, passes, tol, max_cg, seed, yield_stress, adaptive, gravity, speed_clamp,
// density_ratio.
// mode: 0 = f32/f32,
// 1 = f32/f64, 2 = f64.
void* c3_create(const double* cfg, int mode) {
  auto* h = new C3Handle{mode, nullptr};
  orc::C3Config c = c3_cfg_from(cfg);
  if (mode == 2) { auto* s = new orc::C3Sim<double, double>(); s->init(c); h->p = s; }
  else if (mode == 0) { auto* s = new orc::C3Sim<float, float>(); s->init(c); h->p = s; }
  else { auto* s = new orc::C3Sim<float, double>(); s->init(c); h->p = s; }
  return h;
}
void c3_destroy(void* hv) {
  auto* h = (C3Handle*)hv;
  if (!h) return;
  C3DISPATCH(h, delete s);
  delete h;
}
// out: h, L, cw, mass, ncx
void c3_box(const double* cfg, double* out) {
  orc::C3Box b = orc::C3Box::make(c3_cfg_from(cfg));
  out[0] = b.h; out[1] = b.L; out[2] = b.cw; out[3] = b.mass; out[4] = b.ncx;
}
void c3_step(void* hv, int n) {
  auto* h = (C3Handle*)hv;
  for (int k = 0; k < n; k++) C3DISPATCH(h, s->step());
}
size_t c3_count(void* hv) {
  auto* h = (C3Handle*)hv;
  size_t n = 0;
  C3DISPATCH(h, n = s->x.size());
  return n;
}
void c3_state(void* hv, double* x2n, double* v2n) {
  auto* h = (C3Handle*)hv;
  C3DISPATCH(h, {
    for (size_t i = 0; i < s->x.size(); i++) {
      x2n[2 * i] = (double)s->x[i].x; x2n[2 * i + 1] = (double)s->x[i].y;
      v2n[2 * i] = (double)s->v[i].x; v2n[2 * i + 1] = (double)s->v[i].y;
    }
  });
}
void c3_set_state(void* hv, const double* x2n, const double* v2n) {
  auto* h = (C3Handle*)hv;
  C3DISPATCH(h, {
    using RR = decltype(s->x[0].x);
    for (size_t i = 0; i < s->x.size(); i++) {
      s->x[i].x = (RR)x2n[2 * i]; s->x[i].y = (RR)x2n[2 * i + 1];
      s->v[i].x = (RR)v2n[2 * i]; s->v[i].y = (RR)v2n[2 * i + 1];
    }
  });
}
// out: density passes, f64 kinetic energy before the CG, after it and after the speed clamp
void c3_step_stats(void* hv, double* out) {
  auto* h = (C3Handle*)hv;
  C3DISPATCH(h, { out[0] = s->last_passes; out[1] = s->ke_before_cg; out[2] = s->ke_after_cg; out[3] = s->ke_after_clamp; });
}
// out: cg iterations, final residual norm, |b|
void c3_stats(void* hv, double* out) {
  auto* h = (C3Handle*)hv;
  C3DISPATCH(h, { out[0] = s->last_cg; out[1] = (double)s->last_res; out[2] = (double)s->last_bnorm; });
}
double c3_mean_shear(void* hv) {
  auto* h = (C3Handle*)hv;
  double m = 0.0;
  C3DISPATCH(h, m = s->mean_shear());
  return m;
}
void c3_density(void* hv, double* out) {
  auto* h = (C3Handle*)hv;
  C3DISPATCH(h, { for (size_t i = 0; i < s->dens.size(); i++) out[i] = (double)s->dens[i]; });
}
// Neighbour CSR of the last step; returns the edge count (copies only when it fits nb_cap).
int64_t c3_neighbors(void* hv, uint32_t* off, uint32_t* nb, int64_t nb_cap) {
  auto* h = (C3Handle*)hv;
  int64_t e = 0;
  C3DISPATCH(h, {
    e = (int64_t)s->nbr.nb.size();
    std::memcpy(off, s->nbr.off.data(), s->nbr.off.size() * 4);
    if (e <= nb_cap) std::memcpy(nb, s->nbr.nb.data(), (size_t)e * 4);
  });
  return e;
}
// Neighbour CSR over caller points in a box (cfg gives side/dx; L and cells follow from them).
int64_t c3_neighbors_of(const double* cfg, const float* pts, size_t n, uint32_t* off, uint32_t* nb, int64_t nb_cap) {
  orc::C3Box b = orc::C3Box::make(c3_cfg_from(cfg));
  std::vector<orc::V2<float>> p(n);
  for (size_t i = 0; i < n; i++) p[i] = orc::V2<float>(pts[2 * i], pts[2 * i + 1]);
  orc::Nbrs L = orc::c3_neighbors(p, b);
  std::memcpy(off, L.off.data(), L.off.size() * 4);
  if ((int64_t)L.nb.size() <= nb_cap) std::memcpy(nb, L.nb.data(), L.nb.size() * 4);
  return (int64_t)L.nb.size();
}
// Shear rates (f32, pbf.rs:821-849) of caller points with velocities and densities over the given
// neighbour CSR (off: n + 1 entries), or when off is null over the box's lists of those points (cfg
// gives side/dx; h and L follow from them).
void c3_shear_rates_of(const double* cfg, const float* pts, const float* vel, const float* dens, size_t n,
                       float mass, const uint32_t* off, const uint32_t* nb, float* out) {
  orc::C3Box b = orc::C3Box::make(c3_cfg_from(cfg));
  std::vector<orc::V2<float>> p(n), v(n);
  for (size_t i = 0; i < n; i++) {
    p[i] = orc::V2<float>(pts[2 * i], pts[2 * i + 1]);
    v[i] = orc::V2<float>(vel[2 * i], vel[2 * i + 1]);
  }
  orc::Nbrs L;
  if (off) {
    L.off.assign(off, off + n + 1);
    L.nb.assign(nb, nb + off[n]);
  } else {
    L = orc::c3_neighbors(p, b);
  }
  std::vector<float> d(dens, dens + n);
  std::vector<float> r = orc::c3_shear_rates(p, v, L, d, mass, b.h, b);
  std::memcpy(out, r.data(), n * 4);
}
// Bingham effective viscosity (f32, pbf.rs:854-868) of each rate.
void c3_bingham_of(const float* rates, size_t n, float mu, float tau_y, float* out) {
  for (size_t i = 0; i < n; i++) out[i] = orc::c3_bingham(rates[i], mu, tau_y);
}
// Initial state of the workload (f32), for checking the product's seeding.
void c3_seed_state(const double* cfg, float* x2n, float* v2n) {
  orc::C3Config c = c3_cfg_from(cfg);
  orc::C3Box b = orc::C3Box::make(c);
  std::vector<float> x, v;
  orc::c3_seed(c, b, x, v);
  std::memcpy(x2n, x.data(), x.size() * 4);
  std::memcpy(v2n, v.data(), v.size() * 4);
}
}

// ---- stage hooks over caller points (any count) in the box of cfg, for the oracle's own pins.
// mode 0: f32 arithmetic (f32 dots), 2: f64. Points and velocities are given as doubles.
template <class R> static std::vector<orc::V2<R>> c3_pts(const double* p, size_t n) {
  std::vector<orc::V2<R>> out(n);
  for (size_t i = 0; i < n; i++) out[i] = orc::V2<R>((R)p[2 * i], (R)p[2 * i + 1]);
  return out;
}
template <class R>
static void c3_pass_of_t(const double* cfg, const double* pts, size_t n, double* dens, double* lam, double* dp) {
  orc::C3Config c = c3_cfg_from(cfg);
  orc::C3Box b = orc::C3Box::make(c);
  std::vector<orc::V2<R>> x = c3_pts<R>(pts, n);
  orc::Nbrs L = orc::c3_neighbors(x, b);
  std::vector<R> d, l;
  std::vector<orc::V2<R>> p, g;
  orc::c3_density_pass(x, L, b, (R)b.mass, (R)c.rest, d, l, p, g);
  for (size_t i = 0; i < n; i++) {
    dens[i] = (double)d[i];
    lam[i] = (double)l[i];
    dp[2 * i] = (double)p[i].x;
    dp[2 * i + 1] = (double)p[i].y;
  }
}
// Density at the points, Morris weights with mu everywhere, then either the operator applied to
// vel (solve == 0: out = L vel) or the implicit viscosity CG from vel (solve != 0: out = v after
// the CG with cfg's dt, tol and max_cg; stats3 = iterations, residual norm, |b|).
template <class R>
static void c3_visc_of_t(const double* cfg, const double* pts, const double* vel, size_t n, int solve, double* out,
                         double* stats3) {
  orc::C3Config c = c3_cfg_from(cfg);
  orc::C3Box b = orc::C3Box::make(c);
  std::vector<orc::V2<R>> x = c3_pts<R>(pts, n), v = c3_pts<R>(vel, n), o;
  orc::Nbrs L = orc::c3_neighbors(x, b);
  std::vector<R> dens = orc::c3_density_at(x, L, b, (R)b.mass);
  std::vector<R> mu(n, (R)c.mu), W;
  std::vector<orc::V2<R>> D;
  orc::c3_morris(x, L, dens, mu, b, (R)b.mass, W, D);
  if (solve) {
    R res = 0, bn = 0;
    uint32_t it = orc::c3_cg<R, R>(v, L, W, D, (R)c.dt, c.tol, c.max_cg, res, bn);
    o = v;
    if (stats3) { stats3[0] = it; stats3[1] = (double)res; stats3[2] = (double)bn; }
  } else {
    orc::c3_laplacian(L, W, D, v, o);
  }
  for (size_t i = 0; i < n; i++) { out[2 * i] = (double)o[i].x; out[2 * i + 1] = (double)o[i].y; }
}
extern "C" {
// One density pass (pbf.rs:1257-1334) over caller points: density, lambda, dp (2n).
void c3_pass_of(const double* cfg, int mode, const double* pts, size_t n, double* dens, double* lam, double* dp) {
  if (mode == 2) c3_pass_of_t<double>(cfg, pts, n, dens, lam, dp);
  else c3_pass_of_t<float>(cfg, pts, n, dens, lam, dp);
}
void c3_visc_of(const double* cfg, int mode, const double* pts, const double* vel, size_t n, int solve, double* out,
                double* stats3) {
  if (mode == 2) c3_visc_of_t<double>(cfg, pts, vel, n, solve, out, stats3);
  else c3_visc_of_t<float>(cfg, pts, vel, n, solve, out, stats3);
}
// Density at caller points (pbf.rs:2136-2144) and the Morris weights of their edges in list order
// (pbf.rs:312-344, mu everywhere); returns the edge count, copying weights only when they fit.
int64_t c3_weights_of(const double* cfg, const double* pts, size_t n, double* dens, uint32_t* off, uint32_t* nb,
                      double* w, int64_t cap) {
  orc::C3Config c = c3_cfg_from(cfg);
  orc::C3Box b = orc::C3Box::make(c);
  std::vector<orc::V2<double>> x = c3_pts<double>(pts, n);
  orc::Nbrs L = orc::c3_neighbors(x, b);
  std::vector<double> d = orc::c3_density_at(x, L, b, (double)b.mass), mu(n, (double)c.mu), W;
  std::vector<orc::V2<double>> D;
  orc::c3_morris(x, L, d, mu, b, (double)b.mass, W, D);
  std::memcpy(dens, d.data(), n * 8);
  std::memcpy(off, L.off.data(), (n + 1) * 4);
  int64_t e = (int64_t)L.nb.size();
  if (e <= cap) { std::memcpy(nb, L.nb.data(), (size_t)e * 4); std::memcpy(w, W.data(), (size_t)e * 8); }
  return e;
}
}
