// sht_exchange.hip -- the C entry points for the latitude cells of a
// Gaussian grid and the surface exchange grid of two such grids, the plan
// utilities of the coupler's grid generator (tool/gmapgen/gmapgen_main.f90:336-405,
// generate_surface_exchage_grid; the same cells at
// exp/APEI07Couple/ebm/EBM_main.f90:200-204).  The arithmetic is in
// sht_host.cpp (host::cell_edges, host::exchange_grid): O(jm) host work over
// the weights a plan computed on the GPU, a grid-generation step rather than a
// transform.
#include "sht_internal.h"

using namespace sht;

extern "C" {

int sht_lat_cell_edges(const double *lat_weight, int jm, double *edges) {
  if (!lat_weight || !edges || jm < 1) return fail(SHT_ERR_INVALID, "sht_lat_cell_edges: bad arguments");
  host::cell_edges(lat_weight, jm, edges);
  return SHT_OK;
}

int sht_exchange_grid(const double *lat_a, const double *w_a, int jm_a, const double *lat_b,
                      const double *w_b, int jm_b, double *lat, double *w, int *jm_out) {
  if (!lat_a || !w_a || !w_b || !lat || !w || !jm_out || jm_a < 1 || jm_b < 1)
    return fail(SHT_ERR_INVALID, "sht_exchange_grid: bad arguments");
  (void)lat_b;  // B contributes its cell edges only
  *jm_out = host::exchange_grid(lat_a, w_a, jm_a, w_b, jm_b, lat, w);
  return SHT_OK;
}

}  // extern "C"
