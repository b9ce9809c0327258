// Host-side derivation of the standard NFFL fuel catalog in f32.
//
// The reference evaluates its catalog at compile time
// (crates/firelib/src/lib.rs:125-268 via Fuel::standard / Fuel::make,
// crates/firelib/src/firelib.rs:554-979).  Its arithmetic rules are kept so
// the device constants carry the same bits:
//   * +, -, *, / are IEEE f32 operations;
//   * exp and powf are evaluated in f64 and rounded to f32
//     (crates/firelib/src/f32.rs:23-34);
//   * every particle quantity makes the imperial -> SI -> imperial round trip
//     of crates/firelib/src/units.rs:33-115 (ParticleDef::standard stores SI,
//     Particle::make converts back), powi(2) = F*F and powi(3) = F*(F*F);
//   * accumulations run in the reference's order: Particle area weights over
//     the definition order, FuelDef sums over the dead-first sorted order.
// The table is lib.rs:126-266 (NFFL13's 10-h load is 0.0580 there).
#include <cmath>
#include <cstring>

#include "pg_internal.h"

namespace pg {
namespace {

enum PType { kDead = 0, kHerb = 1, kWood = 2 };

struct PartSpec {
  PType type;
  float load;  // lb/ft2
  float savr;  // 1/ft
};
struct FuelSpec {
  float depth;  // ft
  float mext;
  int n;
  PartSpec p[4];
};

const FuelSpec kTable[kNumFuels] = {
    {0.1f, 0.01f, 0, {}},
    {1.0f, 0.12f, 1, {{kDead, 0.0340f, 3500.0f}}},
    {1.0f, 0.15f, 4,
     {{kDead, 0.0920f, 3000.0f}, {kDead, 0.0460f, 109.0f}, {kDead, 0.0230f, 30.0f},
      {kHerb, 0.0230f, 1500.0f}}},
    {2.5f, 0.25f, 1, {{kDead, 0.1380f, 1500.0f}}},
    {6.0f, 0.2f, 4,
     {{kDead, 0.2300f, 2000.0f}, {kDead, 0.1840f, 109.0f}, {kDead, 0.0920f, 30.0f},
      {kWood, 0.2300f, 1500.0f}}},
    {2.0f, 0.2f, 3,
     {{kDead, 0.0460f, 2000.0f}, {kDead, 0.0230f, 109.0f}, {kWood, 0.0920f, 1500.0f}}},
    {2.5f, 0.25f, 3,
     {{kDead, 0.0690f, 1750.0f}, {kDead, 0.1150f, 109.0f}, {kDead, 0.0920f, 30.0f}}},
    {2.5f, 0.40f, 4,
     {{kDead, 0.0520f, 1750.0f}, {kDead, 0.0860f, 109.0f}, {kDead, 0.0690f, 30.0f},
      {kWood, 0.0170f, 1550.0f}}},
    {0.2f, 0.30f, 3,
     {{kDead, 0.0690f, 2000.0f}, {kDead, 0.0460f, 109.0f}, {kDead, 0.1150f, 30.0f}}},
    {0.2f, 0.25f, 3,
     {{kDead, 0.1340f, 2500.0f}, {kDead, 0.0190f, 109.0f}, {kDead, 0.0070f, 30.0f}}},
    {1.0f, 0.25f, 4,
     {{kDead, 0.1380f, 2000.0f}, {kDead, 0.0920f, 109.0f}, {kDead, 0.2300f, 30.0f},
      {kWood, 0.0920f, 1500.0f}}},
    {1.0f, 0.15f, 3,
     {{kDead, 0.0690f, 1500.0f}, {kDead, 0.2070f, 109.0f}, {kDead, 0.2530f, 30.0f}}},
    {2.3f, 0.20f, 3,
     {{kDead, 0.1840f, 1500.0f}, {kDead, 0.6440f, 109.0f}, {kDead, 0.7590f, 30.0f}}},
    {3.0f, 0.25f, 3,
     {{kDead, 0.3220f, 1500.0f}, {kDead, 0.0580f, 109.0f}, {kDead, 1.2880f, 30.0f}}},
};

constexpr float kSmidgen = 1e-6f;
constexpr float kFoot = 0.3048f, kPound = 0.4535924f, kBtu = 1054.35f;

// f64-evaluated transcendental rounded to f32 (SoftFloat exp / powf).
inline float exp_r(float x) { return static_cast<float>(std::exp(static_cast<double>(x))); }
inline float pow_r(float x, float y) {
  return static_cast<float>(std::pow(static_cast<double>(x), static_cast<double>(y)));
}
inline float safe_div(float a, float b) { return b > kSmidgen ? a / b : 0.0f; }

// Unit factors exactly as units.rs builds them.
struct Units {
  float load, dens, savr, heat;
  Units() {
    const float f2 = kFoot * kFoot;
    load = kPound / f2;
    dens = kPound / (kFoot * f2);
    savr = 1.0f / kFoot;
    heat = kBtu / kPound;
  }
};

struct Part {  // a particle after Particle::make
  PType type;
  float load, savr, dens, heat, si_total, si_eff;
  float surface, area_weight, size_weight, sigma_factor;
  int size_class, mclass;
  bool dead() const { return type == kDead; }
};

int size_class_of(float savr) {
  if (savr > 1200.0f) return 0;
  if (savr > 192.0f) return 1;
  if (savr > 96.0f) return 2;
  if (savr > 48.0f) return 3;
  if (savr > 16.0f) return 4;
  return 5;
}

void derive(const FuelSpec &spec, DevFuel *out) {
  std::memset(out, 0, sizeof(*out));
  const Units u;
  const int n = spec.n;
  Part made[kMaxParticles];

  // Imperial -> SI -> imperial for each particle definition.
  for (int i = 0; i < n; ++i) {
    Part &p = made[i];
    p.type = spec.p[i].type;
    p.load = (spec.p[i].load * u.load) / u.load;
    p.savr = (spec.p[i].savr * u.savr) / u.savr;
    p.dens = (32.0f * u.dens) / u.dens;
    p.heat = (8000.0f * u.heat) / u.heat;
    p.si_total = 0.0555f;
    p.si_eff = 0.01f;
    p.surface = safe_div(p.load * p.savr, p.dens);
    p.size_class = size_class_of(p.savr);
    p.sigma_factor = exp_r(safe_div(-138.0f, p.savr));
    if (p.dead())
      p.mclass = p.size_class <= 1 ? 0 : (p.size_class <= 3 ? 1 : 2);
    else
      p.mclass = p.type == kHerb ? 3 : 4;
  }
  // Area weight: own surface over the same-life surface, summed in the
  // definition order (ParticleDef::area_weight).
  for (int i = 0; i < n; ++i) {
    float total = 0.0f;
    for (int k = 0; k < n; ++k) total += made[k].dead() == made[i].dead() ? made[k].surface : 0.0f;
    made[i].area_weight = safe_div(made[i].surface, total);
  }
  // Size-class weight: same life and size class (ParticleDef::size_class_weight).
  for (int i = 0; i < n; ++i) {
    float s = 0.0f;
    for (int k = 0; k < n; ++k) {
      const bool same = made[k].dead() == made[i].dead() && made[k].size_class == made[i].size_class;
      s += same ? made[k].area_weight : 0.0f;
    }
    made[i].size_weight = s;
  }
  // Dead particles first, each group in definition order (Fuel::standard).
  Part sorted[kMaxParticles];
  int ns = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < n; ++i)
      if (made[i].dead() == (pass == 0)) sorted[ns++] = made[i];
  int n_dead = 0;
  while (n_dead < ns && sorted[n_dead].dead()) ++n_dead;

  const float depth = (spec.depth * kFoot) / kFoot;
  float total_area = 0.0f;
  for (int i = 0; i < ns; ++i) total_area += sorted[i].surface;

  // FuelDef life sums; index 0 alive, 1 dead.
  float aw[2] = {0, 0}, savr[2] = {0, 0}, heat[2] = {0, 0}, seff[2] = {0, 0},
        lload[2] = {0, 0}, fine[2] = {0, 0};
  for (int life = 0; life < 2; ++life) {
    for (int i = 0; i < ns; ++i) {
      const Part &p = sorted[i];
      const bool in = p.dead() == (life == 1);
      aw[life] += in ? p.surface / total_area : 0.0f;
      savr[life] += in ? p.area_weight * p.savr : 0.0f;
      heat[life] += in ? p.area_weight * p.heat : 0.0f;
      seff[life] += in ? p.area_weight * p.si_eff : 0.0f;
      lload[life] += in ? (p.size_weight * p.load) * (1.0f - p.si_total) : 0.0f;
      if (in)
        fine[life] += life == 1 ? p.load * p.sigma_factor : p.load * exp_r(-500.0f / p.savr);
      else
        fine[life] += 0.0f;
    }
  }
  const float sigma = aw[0] * savr[0] + aw[1] * savr[1];
  float beta_sum = 0.0f;
  for (int i = 0; i < ns; ++i) beta_sum += sorted[i].load / sorted[i].dens;
  const float beta = beta_sum / depth;
  const float ratio = beta / (3.348f / pow_r(sigma, 0.8189f));
  const float s15 = pow_r(sigma, 1.5f);
  const float gamma_max = s15 / (495.0f + 0.0594f * s15);
  const float aa = 133.0f / pow_r(sigma, 0.7913f);
  const float gamma = (gamma_max * pow_r(ratio, aa)) * exp_r(aa * (1.0f - ratio));
  float rx[2];
  for (int life = 0; life < 2; ++life) {
    float eta_s = 1.0f;
    if (seff[life] > kSmidgen) {
      const float e = 0.174f / pow_r(seff[life], 0.19f);
      eta_s = e < 1.0f ? e : 1.0f;
    }
    rx[life] = ((lload[life] * heat[life]) * eta_s) * gamma;
  }
  const float wind_b = 0.02526f * pow_r(sigma, 0.54f);
  const float c = 7.47f * exp_r(-0.133f * pow_r(sigma, 0.55f));
  const float e = 0.715f * exp_r(-0.000359f * sigma);
  float bulk = 0.0f;
  for (int i = 0; i < ns; ++i) bulk += sorted[i].load / depth;

  out->np = ns;
  out->has_live = n_dead < ns;
  out->live_area_weight = aw[0];
  out->dead_area_weight = aw[1];
  out->live_rx = rx[0];
  out->dead_rx = rx[1];
  out->fine_dead_factor = 1.0f / fine[1];
  out->live_ext_factor = 2.9f * safe_div(fine[1], fine[0]);
  out->bulk_density = bulk;
  out->flux_ratio = exp_r((0.792f + 0.681f * std::sqrt(sigma)) * (beta + 0.1f)) /
                    (192.0f + 0.2595f * sigma);
  out->slope_k = 5.275f * pow_r(beta, -0.3f);
  out->wind_b = wind_b;
  out->wind_b_inv = 1.0f / wind_b;
  out->wind_k = c * pow_r(ratio, -e);
  out->wind_e = pow_r(ratio, e) / c;
  out->mext = spec.mext;
  out->inv_mext = 1.0f / spec.mext;
  out->sigma = sigma;
  out->beta = beta;
  out->residence_time = 384.0f / sigma;
  for (int i = 0; i < ns; ++i) {
    out->p[i].mclass = sorted[i].mclass;
    out->p[i].dead = sorted[i].dead() ? 1 : 0;
    out->p[i].area_weight = sorted[i].area_weight;
    out->p[i].sigma_factor = sorted[i].sigma_factor;
    out->p[i].load = sorted[i].load;
  }
}

}  // namespace

void build_catalog(DevCatalog *out) {
  for (int i = 0; i < kNumFuels; ++i) derive(kTable[i], &out->fuel[i]);
}

}  // namespace pg
