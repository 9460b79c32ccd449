// Operation counts of the oracle (the liboracle_count.so build only, compiled
// with counted.h) -- TEST INFRASTRUCTURE.
//
// oracle_count_slice runs what oracle_time_slice runs -- the forward and the
// reference CPU backward (interior colour adjoint, then boundary sampling)
// over the pixel-samples [i0, i1) of a workload -- and returns the arithmetic
// it evaluated, per phase, with the reference GPU's culling: a scene query
// counts only the groups of its 16x16 tile's list.  The weight image, built
// once over all samples, is not counted.
#include "oracle_backward.h"
#include "oracle_forward.h"

namespace cnt {
bool on = false;
int phase = 0;
uint64_t flops[PHASES] = {}, sfu[PHASES] = {};
}  // namespace cnt

namespace oracle {
void grads_zero(const Scene &sc, bool with_translation, Grads &gr);  // oracle_api.cpp
}
using namespace oracle;

extern "C" {

// out[0..6): forward flops, forward sfu, interior flops, interior sfu,
// boundary flops, boundary sfu (whole slice).
void oracle_count_slice(const dvg_scene_desc *desc, const dvg_render_options *opt, const float *d_render,
                        uint64_t i0, uint64_t i1, double *out) {
    Scene sc = read_scene(desc);
    cnt::on = false;
    Ctx cx;
    make_ctx(sc, cx);
    cx.count_tiles = group_tile_ranges(sc, *opt);
    std::vector<float> wimg = weight_image(sc, *opt);
    Grads gr;
    grads_zero(sc, false, gr);
    uint32_t nsx = std::max(opt->samples_x, 1u), nsy = std::max(opt->samples_y, 1u);
    for (int p = 0; p < cnt::PHASES; ++p) cnt::flops[p] = cnt::sfu[p] = 0;
    cnt::phase = 0;
    cnt::on = true;
    forward_samples(sc, *opt, i0, i1);
    cnt::phase = 1;
    backward_samples(cx, *opt, false, d_render, wimg, gr, i0, i1);
    cnt::phase = 2;
    if (!opt->use_prefiltering) cpu::boundary_sampling(cx, nsx, nsy, opt->seed, d_render, wimg, gr, false, i0, i1);
    cnt::on = false;
    for (int p = 0; p < 3; ++p) {
        out[2 * p] = (double)cnt::flops[p];
        out[2 * p + 1] = (double)cnt::sfu[p];
    }
}

}  // extern "C"
