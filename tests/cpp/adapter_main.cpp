// Driver of the INTEGRATION.md device adapter (extracted into
// adapter_core.hpp by the Makefile).
//   adapter_check host            host-only entry points (no device)
//   adapter_check device OUT.bin  2D NS manufactured problem, 6 x 5 quads, P3:
//                                 residual of the interpolated manufactured
//                                 state through DeviceDisc::form_residual,
//                                 written as raw doubles to OUT.bin
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "adapter_core.hpp"

static ice_mesh_desc mms_mesh() {
    ice_mesh_desc md{};
    md.ndim = 2;
    md.nelem[0] = 6;
    md.nelem[1] = 5;
    md.xmax[0] = md.xmax[1] = 1.0;
    for (int k = 0; k < 4; ++k) md.bctype[k] = ICE_BC_DIRICHLET;
    md.periodic_mode = ICE_PERIODIC_REFERENCE;
    return md;
}

static ice_disc_params mms_params() {
    ice_disc_params p;
    iceicle::gpu::check(ice_disc_params_default(ICE_FLUX_NAVIER_STOKES, &p));
    p.visc_mu = 0.003;
    p.dirichlet[0].kind = ICE_FIELD_NS_MMS_STATE;
    p.source.kind = ICE_FIELD_NS_MMS_SOURCE;
    return p;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    try {
        if (std::strcmp(argv[1], "host") == 0) {
            ice_mesh_desc md = mms_mesh();
            ice_mesh_counts cnt{};
            iceicle::gpu::check(ice_mesh_export(&md, &cnt, nullptr, nullptr, nullptr));
            const ice_disc_params p = mms_params();
            std::printf("host: nelem=%d nface=%d gamma=%.2f version=%d\n", cnt.nelem, cnt.nface, p.gamma,
                        ice_version());
            if (cnt.nelem != 30 || cnt.nface != 6 * 4 + 5 * 5 + 22) return 1;
            std::printf("OK\n");
            return 0;
        }
        if (std::strcmp(argv[1], "device") == 0 && argc == 3) {
            iceicle::gpu::DeviceDisc d(mms_mesh(), 3, mms_params());
            std::printf("device: neq=%d n=%zu\n", d.neq, d.n);
            if (d.neq != 4 || d.n != 30u * 16u * 4u) return 1;
            // nodal values of the manufactured state at the element nodes
            std::vector<double> xn(30 * 16 * 2), u(d.n), r(d.n, std::nan(""));
            iceicle::gpu::check(ice_space_node_coords(d.space, xn.data()));
            ice_field f{};
            f.kind = ICE_FIELD_NS_MMS_STATE;
            iceicle::gpu::check(ice_field_eval_points(&f, 2, 4, 30 * 16, xn.data(), u.data()));
            d.form_residual(u.data(), r.data());
            FILE* fp = std::fopen(argv[2], "wb");
            if (!fp || std::fwrite(r.data(), sizeof(double), r.size(), fp) != r.size()) return 1;
            std::fclose(fp);
            double mx = 0;
            for (double v : r) mx = std::fmax(mx, std::fabs(v));
            std::printf("device: max|R| = %.3e\nOK\n", mx);
            return std::isfinite(mx) ? 0 : 1;
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 2;
}
