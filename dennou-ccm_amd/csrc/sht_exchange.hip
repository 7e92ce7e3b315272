// sht_exchange.hip -- latitude cell edges of a Gaussian grid and the surface
// exchange grid of two such grids, the plan utilities of the coupler's grid
// generator (tool/gmapgen/gmapgen_main.f90:336-405, generate_surface_exchage_grid;
// the same cell edges at exp/APEI07Couple/ebm/EBM_main.f90:200-204).
//
// A latitude's cell is the band whose area equals its Gauss weight:
//   edge_0 = -pi/2,  edge_j = asin(w_j + sin(edge_{j-1})),  edge_jm = +pi/2
// (w_j indexed from 1, latitudes south to north; the construction closes at
// +pi/2 only because the weights sum to 2).  The exchange grid of grids A and
// B merges A's interior edges with B's, sorts them and keeps every band of
// nonzero area (|sin e_j - sin e_{j-1}| > 1e-12) with its mid-latitude and its
// area as weight; two grids with the same number of latitudes give grid A
// itself.  O(jm) host arithmetic over the weights a plan computed on the GPU
// (sht_plan_grid, sht_zonal_grid) -- a grid-generation step, not a transform.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sht_internal.h"

namespace sht {
namespace {

// edges[0..jm]: the recurrence in the order the generator evaluates it, so a
// rounding that accumulates along the sum is the generator's own.
void cell_edges(const double *w, int jm, double *edges) {
  edges[0] = -M_PI / 2.0;
  for (int j = 1; j < jm; ++j) edges[j] = std::asin(w[j - 1] + std::sin(edges[j - 1]));
  edges[jm] = M_PI / 2.0;
}

}  // namespace
}  // namespace sht

using namespace sht;

extern "C" {

int sht_lat_cell_edges(const double *lat_weight, int jm, double *edges) {
  if (!lat_weight || !edges || jm < 1) return fail(SHT_ERR_INVALID, "sht_lat_cell_edges: bad arguments");
  cell_edges(lat_weight, jm, edges);
  return SHT_OK;
}

int sht_exchange_grid(const double *lat_a, const double *w_a, int jm_a, const double *lat_b,
                      const double *w_b, int jm_b, double *lat, double *w, int *jm_out) {
  if (!lat_a || !w_a || !w_b || !lat || !w || !jm_out || jm_a < 1 || jm_b < 1)
    return fail(SHT_ERR_INVALID, "sht_exchange_grid: bad arguments");
  (void)lat_b;  // B contributes its edges only
  if (jm_a == jm_b) {
    // gmapgen_main.f90:349-355: the exchange grid is grid A
    std::copy(lat_a, lat_a + jm_a, lat);
    std::copy(w_a, w_a + jm_a, w);
    *jm_out = jm_a;
    return SHT_OK;
  }
  std::vector<double> ea(jm_a + 1), eb(jm_b + 1);
  cell_edges(w_a, jm_a, ea.data());
  cell_edges(w_b, jm_b, eb.data());
  // A's edges 0..jm_a-1 and B's 1..jm_b (gmapgen_main.f90:372-373)
  std::vector<double> e(ea.begin(), ea.end() - 1);
  e.insert(e.end(), eb.begin() + 1, eb.end());
  std::sort(e.begin(), e.end());
  int n = 0;
  for (size_t j = 1; j < e.size(); ++j) {
    const double area = std::sin(e[j]) - std::sin(e[j - 1]);
    if (std::fabs(area) > 1e-12) {
      lat[n] = 0.5 * (e[j - 1] + e[j]);
      w[n] = area;
      ++n;
    }
  }
  *jm_out = n;
  return SHT_OK;
}

}  // extern "C"
