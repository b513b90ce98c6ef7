// Tracker::optimize for one frame (Tracker.cpp:337-580): particle swarm
// (PSO.cpp:198-295) followed by the downhill simplex polish
// (DownhillSimplex.cpp:39-172), or the simplex alone.
//
// Cost evaluations go through the scene's fused batch path. Two PSO forms:
//   asynchronous (reference): particles are moved and evaluated one at a
//     time and gBest updates immediately, so particle k follows the gBest
//     that particles < k just improved (PSO.cpp:236-256);
//   synchronous: every particle of an epoch moves against the gBest of the
//     previous epoch and the whole epoch is one fused batch; pBest/gBest are
//     then updated in particle order.
// Random numbers reproduce glibc srand()/rand() (PSO.cpp:131-159) through
// random_r on a private TYPE_3 state, so concurrent optimisations do not
// share the process-wide generator.
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "geometry.hpp"


namespace {

using asc::Frame;

class GlibcRand {
 public:
  explicit GlibcRand(unsigned seed)
  {
    std::memset(&data_, 0, sizeof data_);
    std::memset(state_, 0, sizeof state_);
    initstate_r(seed, state_, sizeof state_, &data_);
  }
  int next()
  {
    int32_t r;
    random_r(&data_, &r);
    return r;
  }
  // getRandomClamped, PSO.cpp:156-159
  float clamped() { return (float)next() / (float)RAND_MAX; }
  // getRandom, PSO.cpp:151-154
  float range(float lo, float hi) { return lo + clamped() * (hi - lo); }

 private:
  random_data data_;
  char state_[128];
};

struct Particle {
  float ncc = FLT_MAX;
  float pos[6] = {0, 0, 0, 0, 0, 0};
  float vel[6] = {0, 0, 0, 0, 0, 0};
};

// A cost function of n points of ndim values (n = 1 for the reference's
// one-at-a-time calls, n = particles for a synchronous PSO epoch).
struct Objective {
  asc_cost_fn fn;
  void* user;
  int64_t evals = 0;
  int status = ASC_OK;
  int eval(const double* x, int n, double* costs)
  {
    const int rc = fn(x, n, costs, user);
    if (rc != ASC_OK) {
      status = rc;
      for (int i = 0; i < n; i++) costs[i] = NAN;
    }
    evals += n;
    return rc;
  }
  double one(const double* x)
  {
    double c;
    eval(x, 1, &c);
    return c;
  }
};

// Tracker::minimizationFunc on manipulator offsets relative to the stored
// pose, through the scene's fused batch path.
struct SceneCost {
  asc_scene scene;
  const float* store;
  int volume;
};

int scene_cost(const double* manips, int n, double* costs, void* user)
{
  const SceneCost* sc = (const SceneCost*)user;
  std::vector<float> stores((size_t)n * 7);
  for (int i = 0; i < n; i++) std::memcpy(&stores[(size_t)i * 7], sc->store, 7 * sizeof(float));
  std::vector<int> vols(n, sc->volume);
  return asc_scene_eval_manips(sc->scene, vols.data(), stores.data(), manips, n, costs, nullptr);
}

int pso(Objective& obj, GlibcRand& rng, float lo, float hi, unsigned max_epochs, unsigned max_stall, int np,
        bool synchronous, Particle& g_out)
{
  const float c1 = 1.5f, c2 = 1.5f;  // PSO.hpp:19-20
  std::vector<Particle> p(np);
  for (int i = 1; i < np; i++)
    for (int d = 0; d < 6; d++) p[i].pos[d] = rng.range(lo, hi);
  std::vector<Particle> pb = p;
  Particle g;
  int stall = 0;
  unsigned counter = 0;
  float omega = 0.8f;
  bool go = true;
  std::vector<double> manips((size_t)np * 6), costs(np);
  while (go) {
    if (counter >= max_epochs) go = false;
    const Particle cur = g;
    if (!synchronous) {
      for (int i = 0; i < np; i++) {
        for (int d = 0; d < 6; d++) {
          const float rp = rng.clamped();
          const float rg = rng.clamped();
          p[i].vel[d] = omega * p[i].vel[d] + c1 * rp * (pb[i].pos[d] - p[i].pos[d]) +
                        c2 * rg * (g.pos[d] - p[i].pos[d]);
          p[i].pos[d] += p[i].vel[d];
        }
        double x[6];
        for (int d = 0; d < 6; d++) x[d] = (double)p[i].pos[d];
        p[i].ncc = (float)obj.one(x);
        if (obj.status) return obj.status;
        if (p[i].ncc < pb[i].ncc) pb[i] = p[i];
        if (p[i].ncc < g.ncc) g = p[i];
      }
    } else {
      for (int i = 0; i < np; i++)
        for (int d = 0; d < 6; d++) {
          const float rp = rng.clamped();
          const float rg = rng.clamped();
          p[i].vel[d] = omega * p[i].vel[d] + c1 * rp * (pb[i].pos[d] - p[i].pos[d]) +
                        c2 * rg * (g.pos[d] - p[i].pos[d]);
          p[i].pos[d] += p[i].vel[d];
          manips[(size_t)i * 6 + d] = (double)p[i].pos[d];
        }
      if (obj.eval(manips.data(), np, costs.data())) return obj.status;
      for (int i = 0; i < np; i++) {
        p[i].ncc = (float)costs[i];
        if (p[i].ncc < pb[i].ncc) pb[i] = p[i];
        if (p[i].ncc < g.ncc) g = p[i];
      }
    }
    omega = omega * 0.9f;
    const float diff = std::fabs(g.ncc - cur.ncc);
    if (diff < 1e-4f)
      stall++;
    else if (diff > 0.001f)
      stall = 0;
    if (stall == (int)max_stall) go = false;
    counter++;
  }
  g_out = g;
  return (int)counter;
}

// Nelder-Mead with the centroid-distance stop, DownhillSimplex.cpp:39-172.
// P holds ndim+1 rows of ndim values; Y the costs of the rows.
// speculative: evaluate each iteration's candidate points together (the
// reflection, the expansion, the contraction of the worst vertex and of the
// reflection) and a shrink's vertices together. Costs do not depend on how
// points are grouped, so every decision, and the trajectory, is the one of
// the one-at-a-time order; only the number of evaluations grows.
int amoeba(Objective& obj, std::vector<double>& P, std::vector<double>& Y, int ndim, double ftol,
           bool speculative = false)
{
  const int itmax = 1000;
  const double alpha = 1.0, beta = 0.5, gamma = 2.0;
  const int np = ndim + 1;
  auto at = [&](int i, int j) -> double& { return P[(size_t)i * ndim + j]; };
  std::vector<double> pr(ndim), prr(ndim), pbar(ndim), cent(ndim);
  int iter = 0;
  for (;;) {
    int ilo = 0, ihi, inhi;
    if (Y[0] > Y[1]) {
      ihi = 0;
      inhi = 1;
    } else {
      ihi = 1;
      inhi = 0;
    }
    for (int i = 0; i < np; i++) {
      if (Y[i] < Y[ilo]) ilo = i;
      if (Y[i] > Y[ihi]) {
        inhi = ihi;
        ihi = i;
      } else if (Y[i] > Y[inhi] && i != ihi) {
        inhi = i;
      }
    }
    for (int j = 0; j < ndim; j++) {
      cent[j] = 0;
      for (int i = 0; i < np; i++) cent[j] += at(i, j);
      cent[j] /= np;
    }
    double rtol = 0;
    for (int i = 0; i < np; i++) {
      double ss = 0;
      for (int j = 0; j < ndim; j++) ss += (cent[j] - at(i, j)) * (cent[j] - at(i, j));
      rtol += sqrt(ss);
    }
    rtol /= np;
    if (rtol < ftol || iter == itmax) return iter;
    iter++;
    for (int j = 0; j < ndim; j++) pbar[j] = 0.0;
    for (int i = 0; i < np; i++)
      if (i != ihi)
        for (int j = 0; j < ndim; j++) pbar[j] += at(i, j);
    for (int j = 0; j < ndim; j++) {
      pbar[j] /= 1.0 * ndim;
      pr[j] = (1.0 + alpha) * pbar[j] - alpha * at(ihi, j);
    }
    if (speculative) {
      // candidates: 0 reflection, 1 expansion, 2 contraction of the worst
      // vertex, 3 contraction of the reflection (the worst vertex when the
      // reflection replaced it)
      std::vector<double> cand((size_t)4 * ndim), yc(4);
      for (int j = 0; j < ndim; j++) {
        cand[j] = pr[j];
        cand[ndim + j] = gamma * pr[j] + (1.0 - gamma) * pbar[j];
        cand[2 * ndim + j] = beta * at(ihi, j) + (1.0 - beta) * pbar[j];
        cand[3 * ndim + j] = beta * pr[j] + (1.0 - beta) * pbar[j];
      }
      obj.eval(cand.data(), 4, yc.data());
      if (obj.status) return -1;
      const double ypr = yc[0];
      if (ypr <= Y[ilo]) {
        const int k = yc[1] < Y[ilo] ? 1 : 0;
        for (int j = 0; j < ndim; j++) at(ihi, j) = cand[k * ndim + j];
        Y[ihi] = yc[k];
      } else if (ypr >= Y[inhi]) {
        int k = 2;
        if (ypr < Y[ihi]) {
          for (int j = 0; j < ndim; j++) at(ihi, j) = pr[j];
          Y[ihi] = ypr;
          k = 3;
        }
        if (yc[k] < Y[ihi]) {
          for (int j = 0; j < ndim; j++) at(ihi, j) = cand[k * ndim + j];
          Y[ihi] = yc[k];
        } else {
          std::vector<double> pts((size_t)ndim * ndim), ys(ndim);
          int m = 0;
          for (int i = 0; i < np; i++)
            if (i != ilo) {
              for (int j = 0; j < ndim; j++) {
                at(i, j) = 0.5 * (at(i, j) + at(ilo, j));
                pts[(size_t)m * ndim + j] = at(i, j);
              }
              m++;
            }
          obj.eval(pts.data(), m, ys.data());
          m = 0;
          for (int i = 0; i < np; i++)
            if (i != ilo) Y[i] = ys[m++];
        }
      } else {
        for (int j = 0; j < ndim; j++) at(ihi, j) = pr[j];
        Y[ihi] = ypr;
      }
      if (obj.status) return -1;
      continue;
    }
    const double ypr = obj.one(pr.data());
    if (obj.status) return -1;
    if (ypr <= Y[ilo]) {
      for (int j = 0; j < ndim; j++) prr[j] = gamma * pr[j] + (1.0 - gamma) * pbar[j];
      const double yprr = obj.one(prr.data());
      if (yprr < Y[ilo]) {
        for (int j = 0; j < ndim; j++) at(ihi, j) = prr[j];
        Y[ihi] = yprr;
      } else {
        for (int j = 0; j < ndim; j++) at(ihi, j) = pr[j];
        Y[ihi] = ypr;
      }
    } else if (ypr >= Y[inhi]) {
      if (ypr < Y[ihi]) {
        for (int j = 0; j < ndim; j++) at(ihi, j) = pr[j];
        Y[ihi] = ypr;
      }
      for (int j = 0; j < ndim; j++) prr[j] = beta * at(ihi, j) + (1.0 - beta) * pbar[j];
      const double yprr = obj.one(prr.data());
      if (yprr < Y[ihi]) {
        for (int j = 0; j < ndim; j++) at(ihi, j) = prr[j];
        Y[ihi] = yprr;
      } else {
        // shrink towards the best vertex; each new vertex is evaluated as it
        // is formed, as in the reference
        for (int i = 0; i < np; i++)
          if (i != ilo) {
            for (int j = 0; j < ndim; j++) {
              pr[j] = 0.5 * (at(i, j) + at(ilo, j));
              at(i, j) = pr[j];
            }
            Y[i] = obj.one(pr.data());
          }
      }
    } else {
      for (int j = 0; j < ndim; j++) at(ihi, j) = pr[j];
      Y[ihi] = ypr;
    }
    if (obj.status) return -1;
  }
}

void apply_manip(const double* xyzypr_in, const Frame& pivot, const Frame& manip, double* out)
{
  Frame xc = asc::frame_from_xyzypr(xyzypr_in);
  xc = asc::frame_mul(asc::frame_mul(asc::frame_mul(xc, asc::frame_inverse(pivot)), manip), pivot);
  asc::frame_to_xyzypr(xc, out);
}

}  // namespace

// Tracker::optimize for one frame on any cost of manipulator offsets
// relative to `store` (which the cost reads when it is called, and which is
// updated in place: after the PSO, Tracker.cpp:476-489, and at the end,
// :563-570).
static int optimize_frame_cb(asc_cost_fn fn, void* user, float* store, const Frame& pivot,
                             const asc_optimize_options* opt, asc_optimize_stats* stats)
{
  GlibcRand rng(opt->seed);  // initializeRandom, PSO.cpp:131-149
  Objective obj{fn, user};
  double xyzypr[6];
  asc_pose_store_get(store, xyzypr);
  const double zero[6] = {0, 0, 0, 0, 0, 0};
  const double init_cost = obj.one(zero);
  if (obj.status) return obj.status;
  Frame manip = asc::frame_from_xyz_axis_angle(zero);
  std::vector<double> P(7 * 6), Y(7);
  const bool spec = opt->speculative != 0;
  auto simplex_start = [&]() {
    for (int i = 0; i < 7; i++)
      for (int j = 0; j < 6; j++) P[(size_t)i * 6 + j] = (i == j + 1) ? opt->offsets[j] : 0.0;
    if (spec)
      obj.eval(P.data(), 7, Y.data());
    else
      for (int i = 0; i < 7; i++) Y[i] = obj.one(&P[(size_t)i * 6]);
  };
  asc_optimize_stats st;
  std::memset(&st, 0, sizeof st);
  st.initial_cost = init_cost;
  if (opt->method == 0) {
    Particle g;
    const int epochs = pso(obj, rng, (float)opt->min_limit, (float)opt->max_limit, opt->max_iter, opt->max_stall,
                           opt->particles, opt->synchronous != 0, g);
    if (obj.status) return obj.status;
    double best[6];
    for (int d = 0; d < 6; d++) best[d] = (double)g.pos[d];
    manip = asc::frame_from_xyz_axis_angle(best);
    // Tracker.cpp:476-486: move to the PSO pose and store it (as floats)
    apply_manip(xyzypr, pivot, manip, xyzypr);
    asc_pose_store_set(store, xyzypr);
    st.pso_epochs = epochs;
    st.pso_best_cost = g.ncc;
    simplex_start();
    if (obj.status) return obj.status;
    const int it = amoeba(obj, P, Y, 6, 1e-3, spec);
    if (obj.status) return obj.status;
    const double final_cost = obj.one(&P[0]);
    if (obj.status) return obj.status;
    st.simplex_iterations = it;
    st.final_cost = final_cost;
    manip = final_cost < init_cost ? asc::frame_from_xyz_axis_angle(&P[0]) : asc::frame_from_xyz_axis_angle(zero);
  } else {
    simplex_start();
    if (obj.status) return obj.status;
    const int it = amoeba(obj, P, Y, 6, 1e-3, spec);
    if (obj.status) return obj.status;
    st.simplex_iterations = it;
    st.final_cost = obj.one(&P[0]);
    manip = asc::frame_from_xyz_axis_angle(&P[0]);
  }
  // Tracker.cpp:563-570
  apply_manip(xyzypr, pivot, manip, xyzypr);
  asc_pose_store_set(store, xyzypr);
  st.evaluations = obj.evals;
  if (stats) *stats = st;
  return ASC_OK;
}

extern "C" ASC_API int asc_scene_volume_pivot(asc_scene scene, int volume, double pivot[12]);

extern "C" ASC_API int asc_optimize_frame(asc_scene scene, float store[7], const asc_optimize_options* opt,
                                          asc_optimize_stats* stats)
{
  if (!scene || !store || !opt) return asc::fail(ASC_ERR_INVALID, "bad argument");
  if (opt->particles < 1) return asc::fail(ASC_ERR_INVALID, "particles must be >= 1");
  double pv[12];
  int rc = asc_scene_volume_pivot(scene, opt->volume, pv);
  if (rc) return rc;
  SceneCost sc{scene, store, opt->volume};
  return optimize_frame_cb(scene_cost, &sc, store, asc::frame_load(pv), opt, stats);
}

extern "C" ASC_API int asc_optimize_frame_fn(asc_cost_fn fn, void* user, float store[7], const double pivot[12],
                                             const asc_optimize_options* opt, asc_optimize_stats* stats)
{
  if (!fn || !store || !pivot || !opt) return asc::fail(ASC_ERR_INVALID, "bad argument");
  if (opt->particles < 1) return asc::fail(ASC_ERR_INVALID, "particles must be >= 1");
  return optimize_frame_cb(fn, user, store, asc::frame_load(pivot), opt, stats);
}

extern "C" ASC_API int asc_pso(asc_cost_fn fn, void* user, unsigned seed, float min_limit, float max_limit,
                               unsigned max_epochs, unsigned max_stall, int particles, int synchronous,
                               float best[6], float* best_cost, int* epochs)
{
  if (!fn || !best || particles < 1) return asc::fail(ASC_ERR_INVALID, "asc_pso: bad argument");
  GlibcRand rng(seed);
  Objective obj{fn, user};
  Particle g;
  const int ep = pso(obj, rng, min_limit, max_limit, max_epochs, max_stall, particles, synchronous != 0, g);
  if (obj.status) return obj.status;
  std::memcpy(best, g.pos, sizeof g.pos);
  if (best_cost) *best_cost = g.ncc;
  if (epochs) *epochs = ep;
  return ASC_OK;
}

extern "C" ASC_API int asc_simplex(asc_cost_fn fn, void* user, double* P, double* Y, int ndim, double ftol,
                                   int* iterations)
{
  if (!fn || !P || !Y || ndim < 1 || ndim > 20) return asc::fail(ASC_ERR_INVALID, "asc_simplex: bad argument");
  Objective obj{fn, user};
  std::vector<double> PP(P, P + (size_t)(ndim + 1) * ndim), YY(Y, Y + ndim + 1);
  const int it = amoeba(obj, PP, YY, ndim, ftol);
  if (obj.status) return obj.status;
  std::memcpy(P, PP.data(), PP.size() * sizeof(double));
  std::memcpy(Y, YY.data(), YY.size() * sizeof(double));
  if (iterations) *iterations = it;
  return ASC_OK;
}
