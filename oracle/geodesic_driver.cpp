// Test infrastructure only: a C entry point over the reference's vendored
// exact geodesic library (tfce_mediation/lib/geodesic, the MMP algorithm of
// Surazhsky et al. as implemented by Kirsanov), called the way
// adjacency.pyx:44-79 calls it: for every vertex i, propagate up to the
// largest threshold + 1e-10, then read the distance of every vertex j > i.
// Built by oracle/Makefile into oracle/_ref/libgeoref.so; the product's own
// implementation is tfce_mediation_amd/csrc/geodesic.cpp.
#include <cstdint>
#include <cstring>
#include <vector>

#include "geodesic_algorithm_exact.h"

extern "C" int64_t ref_geodesic_pairs(const float* v, int32_t nv, const int32_t* f, int32_t nf, double max_dist,
                                      int64_t cap, int32_t* pi, int32_t* pj, double* pd) {
  geodesic::Mesh mesh;
  mesh.initialize_mesh_data<float, int32_t>((unsigned)nv, const_cast<float*>(v), (unsigned)nf,
                                            const_cast<int32_t*>(f));
  geodesic::GeodesicAlgorithmExact alg(&mesh);
  std::vector<geodesic::SurfacePoint> src(1);
  const double reach = max_dist + 1e-10;
  int64_t n = 0;
  for (int32_t i = 0; i < nv; ++i) {
    src[0] = geodesic::SurfacePoint(&mesh.vertices()[i]);
    alg.propagate(src, reach);
    for (int32_t j = i + 1; j < nv; ++j) {
      geodesic::SurfacePoint t(&mesh.vertices()[j]);
      double d = 0.0;
      alg.best_source(t, d);
      if (d < max_dist) {
        if (n < cap) {
          pi[n] = i;
          pj[n] = j;
          pd[n] = d;
        }
        ++n;
      }
    }
  }
  return n;
}
