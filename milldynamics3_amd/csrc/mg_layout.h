// Device-side data layout of one drum slot, shared by the host runtime and the HIP kernels.
//
// Every drum of a batch owns one MgDrum record (constants + device pointers into one arena).
// Arrays are structure-of-arrays, float2 for 2D vectors. The host fills the constants from
// Params with the same f32 expressions the reference evaluates (crates/mill-core/src/params.rs,
// pbf.rs, dem.rs), so transcendental-derived constants (lifter and boundary-point trigonometry,
// squeeze-film prefactors, EMA factors) are bit-identical to a CPU evaluation.
#pragma once
#include <stdint.h>

#define MG_MAX_LIFTERS 32
#define MG_MAX_BP 512
#define MG_BLOCK 1024
#define MG_NO_SLOT 0xFFFFFFFFu

// Error bits reported per drum (MgScalars::err).
#define MG_ERR_EDGES 1u        // neighbour-list capacity exceeded
#define MG_ERR_SPARSE 2u       // bounding box needs the sparse grid layout and it overflowed
#define MG_ERR_PAIRS 4u        // ball candidate-pair capacity exceeded
#define MG_ERR_GROUPS 8u       // ball boundary-group capacity exceeded
#define MG_ERR_LINKS 16u       // squeeze-link capacity exceeded
#define MG_ERR_ITEMS 32u       // per-ball item list capacity exceeded
#define MG_ERR_RECS 64u        // ordered-record capacity exceeded
#define MG_ERR_IMPACTS 128u    // impact-record capacity exceeded
#define MG_ERR_TABLE 256u      // stress debounce table full

// Per-substep inputs that depend only on the drum's deterministic angle/time schedule. The host
// evaluates them with the C library so they match a CPU run bit for bit.
struct MgSubstepIn {
  float dt;
  float angle;       // drum angle at the start of the substep
  float angle_next;  // angle + omega * dt (not wrapped), lib.rs / pbf.rs:1236
  float ca0, sa0;    // from_angle(-angle)
  float cn, sn;      // from_angle(-angle_next)
  float cp, sp;      // from_angle(angle_next)
  float alpha;       // 1 - exp(-dt / tau), lib.rs:523
  double now;        // sim_time before the substep (stress debounce clock)
  uint32_t force_density;  // 0 = adaptive; else run exactly this many density passes
  uint32_t force_cg;       // 0 = tolerance-driven; else stop after this many CG iterations
  uint32_t pad[2];
};

// Device-resident per-drum accumulators and last-substep statistics.
struct MgScalars {
  float drum_angle;
  float power_draw_w, torque_nm, collision_rate_per_s, dissipated_power_w;
  float max_disp;
  uint32_t clamp_hits_total;
  uint32_t err;
  uint32_t has_guess;
  // last substep
  uint32_t visc_iters, density_iters, n_groups, n_links, links_in_range, n_pairs, collisions;
  uint32_t n_impacts, n_contacts;
  float mean_shear, fluid_wall_work, dem_wall_work, dem_dissipated, dem_coupling_work;
  float ke_noslip, ke_drag, ke_buoy, ke_visc, ke_clamp;
  double wall_imp_x, wall_imp_y, wall_ang;
  float wall_bins[32];
  uint32_t dem_hist[12];
  // energy budget (lib.rs:61-78)
  double eb_elapsed, eb_shaft, eb_dem_wall, eb_fluid_wall, eb_contact_diss, eb_wall_slip, eb_viscous,
      eb_clamp_removed, eb_interface, eb_fluid_wall_ang;
  // stress accumulator (stress.rs:83-99)
  double st_count[4], st_se[4], st_hist[28], st_above[6], st_cum_wall, st_cum_above[6];
  uint64_t st_updates;
  uint32_t st_table_used;
  uint32_t pad0;
  double sim_time;
};

struct MgImpact {
  uint32_t kind, ball, partner;
  float vn, slip, e;
};

// View of a built uniform grid (dense or sparse CSR); persisted between stage kernels.
struct GridView {
  const uint32_t* start;
  const uint32_t* idx;
  const uint64_t* skeys;  // sparse keys (sorted), or null for dense
  int32_t min_cx, min_cy;
  int64_t W, H;
  int32_t ncells;  // dense: W*H; sparse: number of occupied cells; 0: empty grid
  int32_t dense;
  float cell;
};

// Per-drum values carried from one stage kernel of a substep to the next.
struct MgStage {
  GridView gball, gA, gB, gC;
  uint32_t ngroups, nlinks, in_range, npairs;
  double ke_mark, e_pred, cw;
  float maxd, maxs;
  double hist[28];  // stress histogram of the current substep's impacts
};

struct MgDrum {
  // ---- sizes and capacities
  int32_t nf, nb;
  int32_t e_cap, fcell_cap, bcell_cap, pair_cap, grp_cap, link_cap, item_cap, rec_cap, imp_cap, tbl_cap;
  // ---- constants
  float R, omega, h, fmass, rest, br, bmass, binertia, btrue;
  uint32_t pbf_iters, dem_iters;
  float surface_tension, wettability, wall_no_slip, ball_no_slip, viscosity, yield_stress;
  int32_t rheology;
  float rest_bb, rest_bw, fric_bb, fric_bw, fric_bb_s, fric_bw_s, fric_vscale, rolling;
  uint32_t lcount;
  float l_height, l_base_half, l_top_half;
  float lsin[MG_MAX_LIFTERS], lcos[MG_MAX_LIFTERS];
  uint32_t bp_cnt;                 // boundary points per ball (pbf.rs:216-219) at dx = h/2
  float bp_x[MG_MAX_BP], bp_y[MG_MAX_BP];  // offsets cos/sin(TAU k / cnt) * radius
  // kernel constants evaluated with the reference's f32 expressions
  float k_poly6;      // 4 / (PI h^8)
  float k_spiky;      // -30 / (PI h^5)
  float k_coh;        // 35840 / (209 PI h^8)
  float h6_64;        // h^6 / 64
  float h_adh, k_adh; // min(h, 2r) and 35840 / (163 PI h_adh^8)
  float h_c, k_poly6_hc, w_hc0;  // drag kernel radius r + h, its Poly6 constant, max(poly6(0), 1e-12)
  float v_max;        // fluid speed clamp (pbf.rs:976-982)
  float impact_thr;   // dem.rs:43-46
  double sq_pair[4], sq_wall[4];  // squeeze film r, c1, c2, prefactor (lubrication.rs:50-63)
  double dxd;         // fluid spacing h/2 as f64
  float reach;        // (H_START * dx) as f32
  double m3, scale_j, dd, window;  // stress scales (lib.rs:219-225, stress.rs:118)
  // ---- fluid state
  float2 *fx, *fv, *fguess;
  float* dye;
  // ---- fluid scratch
  float2 *x0, *dp, *excess, *coh, *adh, *lv;
  float *dens, *lam, *mu_eff, *ffac;
  uint32_t* touched;
  uint32_t *gA_start, *gA_idx, *gA_key;  // step-2 grid (cell h), kept through buoyancy
  uint32_t *gB_start, *gB_idx, *gB_key;  // drag grid (cell h + r), then grid_now (cell h)
  int32_t* gA_meta;                      // min_cx, min_cy, W, H, dense, ncells
  int32_t* gB_meta;
  uint64_t *gA_skey, *gB_skey;           // sparse-layout keys (pow2 >= nf entries)
  uint32_t *gA_sidx, *gB_sidx;
  uint32_t* cursor;                      // scratch >= max(fcell_cap, nf, nb, grp/link caps) + 1
  int32_t fpow2, bpow2;
  uint32_t *noff, *nbr;
  float* W;
  float2* D;
  float2 *cg_x, *cg_r, *cg_z, *cg_p, *cg_ap;
  float* cg_inv;
  // ---- ball state
  float2 *bx, *bv;
  float *bth, *bom;
  // ---- ball scratch
  float2 *bx0, *bvpre, *imp, *bv0;
  float *bth0, *bompre, *ang, *toi, *lam_bw, *bw, *bw0;
  uint32_t *gC_start, *gC_idx, *gC_key;  // ball grids (overlap push, links, DEM broadphase)
  int32_t* gC_meta;
  uint64_t* gC_skey;
  uint32_t* gC_sidx;
  uint32_t *pair_i, *pair_j, *ct_perm, *ct_start;
  float* lam_bb;
  uint32_t *g_fluid, *g_ball, *g_bstart, *gf_start, *gf_perm;
  float *g_c, *g_l2;
  float2* g_lc;
  uint32_t *l_ball, *l_other, *l_bstart, *li_start, *li_perm;
  float2* l_n;
  float* l_c;
  // per-ball item scratch (item_cap entries per ball)
  uint32_t* it_j;
  float2* it_v;
  float* it_w;
  float* it_w2;
  uint32_t* it_n;
  // ordered-record scratch
  uint32_t *rec_key, *rec_ord, *rec_perm, *rec_seg;
  float *rec_v0, *rec_v1, *rec_v2;
  // impacts + stress debounce table
  MgImpact* impacts;
  uint64_t* tbl_key;
  double* tbl_time;
  // accumulators
  MgScalars* sc;
  MgStage* st;
  // per-substep schedule (array of MgSubstepIn, indexed by the launch's substep slot)
  const MgSubstepIn* sched;
};
