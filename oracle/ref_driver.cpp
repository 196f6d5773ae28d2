// Test-infrastructure driver that instantiates the reference's own TFCE
// template (tfce_mediation/lib/fast_tfce.hpp, included from the read-only
// reference tree at build time, never copied) behind a C function, the way
// tfce_mediation/tfce.pyx:21-45 binds it.  <cmath> comes first because the
// Cython-generated module reaches fast_tfce.hpp after Python.h, which pulls in
// <cmath>: pow(float, float) then resolves to the float overload.
#include <cmath>
#include <cstdint>
#include <vector>
#include "fast_tfce.hpp"

extern "C" int ref_tfce(float H, float E, float minT, float deltaT, int32_t V,
                        const int64_t *row_ptr, const int32_t *col,
                        const float *image, float *enhn) {
    if (V <= 0) return 0;
    // The reference never terminates when max(image) >= minT and the step is
    // not positive; refuse those inputs instead of hanging the caller.
    float mx = image[0];
    for (int32_t i = 1; i < V; ++i) mx = image[i] > mx ? image[i] : mx;
    float dT = deltaT == 0.0f ? mx / 100 : deltaT;
    if (mx >= minT && !(dT > 0.0f)) return -2;
    std::vector<std::vector<int> > adj(V);
    for (int32_t v = 0; v < V; ++v)
        for (int64_t e = row_ptr[v]; e < row_ptr[v + 1]; ++e) adj[v].push_back(col[e]);
    tfce<float>(H, E, minT, deltaT, adj, image, enhn);
    return 0;
}
