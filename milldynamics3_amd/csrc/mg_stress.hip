// StressAccum on the GPU for a batch of drums (crates/mill-core/src/stress.rs:83-254).
//
// Each drum keeps the accumulator's 49 f64 values in HBM (count, se_hat, hist and above rates,
// cumulative wall work, cumulative energy above each threshold) and an open-addressing table of
// the last counted time per ball pair / ball-wall key (the reference's BTreeMap last_counted).
// One update folds one sub-step's impact records of every drum: one lane per drum walks its
// records in order, because each record's debounce decision depends on the ones before it, and
// every f64 sum is formed in record order as the reference forms it. Per-record work is a probe of
// the table and a few f64 operations; a sub-step holds tens of records per drum.
//
// Bins: se_bin (stress.rs:58-64) is floor((log10(e) + 3) * 4) clamped to [0, 27] with the host
// libm's log10. The bin rises at 27 thresholds; the host finds each one once (the smallest double
// whose bin reaches k, by bisection over the ordered doubles with the same log10) and the device
// counts the thresholds at or below e, which gives the host's bins exactly (log10 is monotone).
//
// Debounce table: linear probing over a power-of-two table per drum. An entry older than the
// window (now - t >= window) reads as absent for every later update (the clock does not go back:
// a decreasing now_s is rejected), so its slot is reused by a new key found absent; this replaces
// the reference's periodic retain (stress.rs:122-124), which only drops such entries. A table with
// neither an empty nor a reusable slot on a key's probe path sets the drum's error flag.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/millgpu.h"
#include "mg_internal.h"

#define ST_VALUES 49
#define ST_LANES 64
#define ST_BINS 28
#define ST_EMPTY 0xFFFFFFFFFFFFFFFFull

namespace {

struct StDrum {  // per drum, per update
  double m3, scale_j, now, sub_dt, alpha, wall_work, window;
};

__device__ __forceinline__ uint32_t st_slot(uint64_t key, uint32_t bits) {
  return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

}  // namespace

// One lane per drum. acc[d * 49 ..]: count 4, se 4, hist 28, above 6, cum_wall 1, cum_above 6.
extern "C" __global__ void __launch_bounds__(ST_LANES)
st_update(const mg_impact_record* __restrict__ rec, const uint32_t* __restrict__ beg,
          const uint32_t* __restrict__ end, const StDrum* __restrict__ par,
          const double* __restrict__ thr_bin, uint32_t n_drums, double* __restrict__ acc, uint64_t* __restrict__ keys,
          double* __restrict__ times, uint32_t cap_bits, uint32_t* __restrict__ err) {
  __shared__ double s_hist[ST_LANES][ST_BINS];
  __shared__ double s_cls[ST_LANES][8];  // count 4, se_hat 4 of this update (indexed by class)
  const uint32_t t = threadIdx.x, d = blockIdx.x * ST_LANES + t;
  for (int b = 0; b < ST_BINS; b++) s_hist[t][b] = 0.0;
  for (int c = 0; c < 8; c++) s_cls[t][c] = 0.0;
  if (d >= n_drums) return;
  const StDrum p = par[d];
  const double THR[6] = {0.03, 0.1, 0.3, 1.0, 3.0, 10.0};
  const uint32_t cap = 1u << cap_bits;
  uint64_t* K = keys + (size_t)d * cap;
  double* T = times + (size_t)d * cap;
  double above[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  uint32_t full = 0;
  for (uint32_t r = beg[d]; r < end[d]; r++) {
    const mg_impact_record q = rec[r];
    if (p.window > 0.0) {
      const uint32_t a = q.partner == 0xFFFFFFFFu ? q.ball : min(q.ball, q.partner);
      const uint32_t b = q.partner == 0xFFFFFFFFu ? 0xFFFFFFFFu : max(q.ball, q.partner);
      const uint64_t key = ((uint64_t)a << 32) | b;
      uint32_t s = st_slot(key, cap_bits), reuse = 0xFFFFFFFFu, stop = 0xFFFFFFFFu;
      bool found = false;
      for (uint32_t step = 0; step < cap; step++, s = (s + 1) & (cap - 1)) {
        const uint64_t k = K[s];
        if (k == ST_EMPTY) { stop = s; break; }
        if (k == key) { found = true; stop = s; break; }
        if (reuse == 0xFFFFFFFFu && !(p.now - T[s] < p.window)) reuse = s;
      }
      if (found) {
        if (p.now - T[stop] < p.window) continue;  // the same episode
        T[stop] = p.now;
      } else {
        const uint32_t at = reuse != 0xFFFFFFFFu ? reuse : stop;
        if (at == 0xFFFFFFFFu) {
          full = 1;  // no room: counted as the reference would, the key is not remembered
        } else {
          K[at] = key;
          T[at] = p.now;
        }
      }
    }
    // classify (stress.rs:68-79)
    const bool bb = q.kind == 0;
    const double vn = (double)q.normal_speed, vs = (double)q.slip_speed;
    const int cls = bb ? (vs > vn ? 1 : 0) : (vs > vn ? 3 : 2);
    const double m_red = bb ? 0.5 * p.m3 : p.m3;
    const double se = 0.5 * m_red * (vn * vn + vs * vs);
    const double eh = se / p.scale_j;
    s_cls[t][cls] += 1.0;
    s_cls[t][4 + cls] += eh;
    int bin = 0;
    if (eh > 0.0)  // NaN and non-positive energies land in bin 0
      for (int k = 0; k < ST_BINS - 1; k++) bin += eh >= thr_bin[k] ? 1 : 0;
    s_hist[t][bin] += 1.0;
    for (int k = 0; k < 6; k++)
      if (eh >= THR[k]) above[k] += se;
  }
  double* A = acc + (size_t)d * ST_VALUES;
  const double a = p.alpha, dt = p.sub_dt;
  for (int c = 0; c < 4; c++) {
    A[c] += a * (s_cls[t][c] / dt - A[c]);
    A[4 + c] += a * (s_cls[t][4 + c] / dt - A[4 + c]);
  }
  for (int b = 0; b < ST_BINS; b++) A[8 + b] += a * (s_hist[t][b] / dt - A[8 + b]);
  for (int k = 0; k < 6; k++) A[36 + k] += a * (above[k] / dt - A[36 + k]);
  A[42] += p.wall_work;
  for (int k = 0; k < 6; k++) A[43 + k] += above[k];
  if (full) err[d] = 1;
}

struct mg_stress {
  uint32_t n = 0, cap_bits = 0;
  char* dev = nullptr;  // acc | keys | times | thresholds | err | params | offsets scratch
  double* acc = nullptr;
  uint64_t* keys = nullptr;
  double* times = nullptr;
  double* thr = nullptr;
  uint32_t* err = nullptr;
  StDrum* par = nullptr;
  std::vector<double> factor, last_now;
  std::vector<uint8_t> started;
  std::string msg;
};

namespace {

// stress.rs:58-64 with the host libm's log10 (the reference's f64::log10).
uint32_t se_bin_host(double e) {
  if (std::isnan(e) || e <= 0.0) return 0;
  const double t = (std::log10(e) - -3.0) * 4.0;
  const double c = std::fmax(t, 0.0);
  const uint64_t u = c >= 1.8446744073709552e19 ? UINT64_MAX : (uint64_t)c;
  return (uint32_t)std::min<uint64_t>(u, 27);
}

// thr[k - 1] = the smallest positive double whose bin is at least k (k = 1..27).
void bin_thresholds(double* thr) {
  for (uint32_t k = 1; k < ST_BINS; k++) {
    uint64_t lo = 1, hi = 0x7FF0000000000000ull;  // bin(lo) < k <= bin(hi = inf)
    while (hi - lo > 1) {
      const uint64_t mid = lo + (hi - lo) / 2;
      double v;
      std::memcpy(&v, &mid, 8);
      if (se_bin_host(v) >= k) hi = mid; else lo = mid;
    }
    std::memcpy(&thr[k - 1], &hi, 8);
  }
}

int set_msg(mg_stress* h, char* err, size_t errlen, int rc, const std::string& m) {
  if (h) h->msg = m;
  if (err && errlen) {
    std::strncpy(err, m.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return rc;
}

}  // namespace

extern "C" {

int mg_stress_create(uint32_t n_drums, uint32_t table_capacity, mg_stress** out, char* err, size_t errlen) {
  if (!out || n_drums == 0) return set_msg(nullptr, err, errlen, MG_E_ARG, "n_drums must be positive");
  uint32_t bits = 6;
  while ((1u << bits) < table_capacity && bits < 24) bits++;
  if ((1u << bits) < table_capacity) return set_msg(nullptr, err, errlen, MG_E_ARG, "table_capacity above 2^24");
  auto* h = new mg_stress();
  h->n = n_drums;
  h->cap_bits = bits;
  const size_t cap = (size_t)1 << bits;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_acc = 0, o_keys = o_acc + up((size_t)n_drums * ST_VALUES * 8);
  const size_t o_times = o_keys + up((size_t)n_drums * cap * 8);
  const size_t o_thr = o_times + up((size_t)n_drums * cap * 8);
  const size_t o_err = o_thr + up(ST_BINS * 8);
  const size_t o_par = o_err + up((size_t)n_drums * 4);
  const size_t total = o_par + up((size_t)n_drums * sizeof(StDrum));
  if (hipMalloc(&h->dev, total) != hipSuccess) {
    delete h;
    return set_msg(nullptr, err, errlen, MG_E_DEVICE, "hipMalloc failed for the stress accumulators");
  }
  h->acc = reinterpret_cast<double*>(h->dev + o_acc);
  h->keys = reinterpret_cast<uint64_t*>(h->dev + o_keys);
  h->times = reinterpret_cast<double*>(h->dev + o_times);
  h->thr = reinterpret_cast<double*>(h->dev + o_thr);
  h->err = reinterpret_cast<uint32_t*>(h->dev + o_err);
  h->par = reinterpret_cast<StDrum*>(h->dev + o_par);
  double thr[ST_BINS - 1];
  bin_thresholds(thr);
  const bool ok = hipMemset(h->acc, 0, (size_t)n_drums * ST_VALUES * 8) == hipSuccess &&
                  hipMemset(h->keys, 0xFF, (size_t)n_drums * cap * 8) == hipSuccess &&
                  hipMemset(h->err, 0, (size_t)n_drums * 4) == hipSuccess &&
                  hipMemcpy(h->thr, thr, sizeof thr, hipMemcpyHostToDevice) == hipSuccess &&
                  hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    (void)hipFree(h->dev);
    delete h;
    return set_msg(nullptr, err, errlen, MG_E_DEVICE, "HIP error initialising the stress accumulators");
  }
  h->factor.assign(n_drums, 0.25);  // DEBOUNCE_FACTOR (stress.rs:49)
  h->last_now.assign(n_drums, 0.0);
  h->started.assign(n_drums, 0);
  *out = h;
  return MG_OK;
}

void mg_stress_destroy(mg_stress* h) {
  if (!h) return;
  if (h->dev) (void)hipFree(h->dev);
  delete h;
}

int mg_stress_set_debounce_factor(mg_stress* h, uint32_t drum, double factor) {
  if (!h || drum >= h->n || !std::isfinite(factor)) return MG_E_ARG;
  h->factor[drum] = factor;
  return MG_OK;
}

int mg_stress_update(mg_stress* h, const mg_impact_record* d_records, const uint32_t* d_offsets,
                     const mg_stress_step* steps, char* err, size_t errlen) {
  if (!h || !d_offsets || !steps) return set_msg(h, err, errlen, MG_E_ARG, "null argument");
  return mg_stress_update_ranges(h, d_records, d_offsets, d_offsets + 1, steps, err, errlen);
}

int mg_stress_update_ranges(mg_stress* h, const mg_impact_record* d_records, const uint32_t* d_beg,
                            const uint32_t* d_end, const mg_stress_step* steps, char* err, size_t errlen) {
  if (!h || !d_beg || !d_end || !steps) return set_msg(h, err, errlen, MG_E_ARG, "null argument");
  std::vector<StDrum> par(h->n);
  for (uint32_t d = 0; d < h->n; d++) {
    const mg_stress_step& s = steps[d];
    if (h->started[d] && s.now_s < h->last_now[d])
      return set_msg(h, err, errlen, MG_E_ARG, "drum " + std::to_string(d) + ": now_s went back in time");
    par[d] = StDrum{s.m3, s.scale_j, s.now_s, s.sub_dt, s.alpha, s.wall_work_j,
                    h->factor[d] * std::sqrt(s.d / 9.81)};  // stress.rs:121
  }
  if (hipMemcpy(h->par, par.data(), par.size() * sizeof(StDrum), hipMemcpyHostToDevice) != hipSuccess)
    return set_msg(h, err, errlen, MG_E_DEVICE, "HIP copy error");
  hipLaunchKernelGGL(st_update, dim3((h->n + ST_LANES - 1) / ST_LANES), dim3(ST_LANES), 0, nullptr, d_records, d_beg,
                     d_end, h->par, h->thr, h->n, h->acc, h->keys, h->times, h->cap_bits, h->err);
  std::vector<uint32_t> e(h->n);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(e.data(), h->err, e.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return set_msg(h, err, errlen, MG_E_DEVICE, "HIP error in the stress update");
  for (uint32_t d = 0; d < h->n; d++) {
    h->started[d] = 1;
    h->last_now[d] = steps[d].now_s;
  }
  for (uint32_t d = 0; d < h->n; d++)
    if (e[d])
      return set_msg(h, err, errlen, MG_E_CAPACITY,
                     "drum " + std::to_string(d) + ": debounce table full (create with a larger table_capacity)");
  return MG_OK;
}

size_t mg_stress_blob_bytes(const mg_stress* h) {
  return h ? 8 * ST_VALUES + 16 * ((size_t)1 << h->cap_bits) + 24 : 0;
}

int mg_stress_save(mg_stress* h, uint32_t drum, void* out) {
  if (!h || drum >= h->n || !out) return MG_E_ARG;
  const size_t cap = (size_t)1 << h->cap_bits;
  char* p = static_cast<char*>(out);
  bool ok = hipMemcpy(p, h->acc + (size_t)drum * ST_VALUES, 8 * ST_VALUES, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(p + 8 * ST_VALUES, h->keys + drum * cap, 8 * cap, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(p + 8 * ST_VALUES + 8 * cap, h->times + drum * cap, 8 * cap, hipMemcpyDeviceToHost) == hipSuccess;
  p += 8 * ST_VALUES + 16 * cap;
  const double meta[3] = {h->factor[drum], h->last_now[drum], (double)h->started[drum]};
  std::memcpy(p, meta, 24);
  return ok ? MG_OK : MG_E_DEVICE;
}

int mg_stress_load(mg_stress* h, uint32_t drum, const void* in) {
  if (!h || drum >= h->n || !in) return MG_E_ARG;
  const size_t cap = (size_t)1 << h->cap_bits;
  const char* p = static_cast<const char*>(in);
  bool ok = hipMemcpy(h->acc + (size_t)drum * ST_VALUES, p, 8 * ST_VALUES, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->keys + drum * cap, p + 8 * ST_VALUES, 8 * cap, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(h->times + drum * cap, p + 8 * ST_VALUES + 8 * cap, 8 * cap, hipMemcpyHostToDevice) == hipSuccess;
  p += 8 * ST_VALUES + 16 * cap;
  double meta[3];
  std::memcpy(meta, p, 24);
  h->factor[drum] = meta[0];
  h->last_now[drum] = meta[1];
  h->started[drum] = meta[2] != 0.0 ? 1 : 0;
  return ok ? MG_OK : MG_E_DEVICE;
}

int mg_stress_reset(mg_stress* h, uint32_t drum) {
  if (!h || drum >= h->n) return MG_E_ARG;
  const size_t cap = (size_t)1 << h->cap_bits;
  h->last_now[drum] = 0.0;
  h->started[drum] = 0;
  return hipMemset(h->acc + (size_t)drum * ST_VALUES, 0, 8 * ST_VALUES) == hipSuccess &&
                 hipMemset(h->keys + drum * cap, 0xFF, 8 * cap) == hipSuccess
             ? MG_OK
             : MG_E_DEVICE;
}

int mg_stress_state(mg_stress* h, uint32_t drum, double* out49) {
  if (!h || drum >= h->n || !out49) return MG_E_ARG;
  return hipMemcpy(out49, h->acc + (size_t)drum * ST_VALUES, ST_VALUES * 8, hipMemcpyDeviceToHost) == hipSuccess
             ? MG_OK
             : MG_E_DEVICE;
}

int mg_stress_get_metrics(mg_stress* h, uint32_t drum, uint32_t n_balls, double m3, double scale_j, double d,
                      double power_draw_w, mg_stress_metrics* out) {
  if (!out) return MG_E_ARG;
  double a[ST_VALUES];
  const int rc = mg_stress_state(h, drum, a);
  if (rc) return rc;
  const double* cr = a;
  const double* se = a + 4;
  const double n = (double)std::max<uint32_t>(n_balls, 1);
  mg_stress_metrics& M = *out;
  M.ball_mass_3d_kg = (float)m3;
  M.scale_energy_j = (float)scale_j;
  const double total = cr[0] + cr[1] + cr[2] + cr[3];
  for (int c = 0; c < 4; c++) {
    M.impacts_per_ball_per_s[c] = (float)(cr[c] / n);
    M.mean_se_scaled[c] = cr[c] > 0.0 ? (float)(se[c] / cr[c]) : 0.0f;
  }
  M.impacts_total_per_ball_per_s = (float)(total / n);
  auto pooled = [&](int x, int y) {
    const double w = cr[x] + cr[y];
    return w > 0.0 ? (float)((se[x] + se[y]) / w) : 0.0f;
  };
  M.mean_se_scaled_drop = pooled(0, 2);
  M.mean_se_scaled_shear = pooled(1, 3);
  M.stress_intensity_scaled = total > 0.0 ? (float)((se[0] + se[1] + se[2] + se[3]) / total) : 0.0f;
  const double ppb = power_draw_w * d / n;
  M.wall_power_per_ball_w = (float)ppb;
  M.specific_power_w_per_kg = (float)(ppb / m3);
  const double THR[6] = {0.03, 0.1, 0.3, 1.0, 3.0, 10.0};
  for (int k = 0; k < 6; k++) {
    M.thresholds_scaled[k] = (float)THR[k];
    M.effective_power_per_ball_w[k] = (float)(a[36 + k] / n);
    M.cumulative_effective_energy_per_ball_j[k] = a[43 + k] / n;
  }
  M.cumulative_wall_energy_per_ball_j = a[42] * d / n;
  M.cumulative_specific_energy_j_per_kg = a[42] * d / n / m3;
  return MG_OK;
}

}  // extern "C"
