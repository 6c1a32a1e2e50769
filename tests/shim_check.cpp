// shim_check -- drives the reference-shaped C++ drop-ins of host/gfm2d.hpp on
// host grids, for tests/test_gpu_shims.py, which compares what they return
// with the oracle.
//
//   shim_check <in.bin> <out.bin>
//
// in.bin is float64 throughout: a 12-value header
//   mode Nx Ny numGC x0 y0 dx dy BC_L BC_R BC_B BC_T   (BC: 0 transmissive, 1 reflective, 2 periodic)
// then, by mode,
//   1 FS:      scheme(0 MUSCL, 1 Godunov) gamma pinf dt imin imax jmin jmax, grid
//              -> grid after FS_*::pure_fluid_update, future_grid
//   2 ghost:   variant ext slip gamma1 pinf1 gamma2 pinf2 t, phi, grid1, grid2
//              -> grid1, grid2, realcells1 (4), realcells2 (4), the stored velocity field
//   3 extrap:  N, phi, prims1, prims2           -> prims1, prims2 after extrapolate_vector
//   4 BCs:     grid                             -> grid after apply_BCs_euler
//   5 level set: order every iters t dt, phi, velocity field
//              -> phi after Levelset::update_interface_coupled
//   6 pure RS: paralleltox gamma pinf n, L (n x 4), R (n x 4)   -> flux (n x 4)
//   7 mixed RS: gL pL gR pR n, L, R             -> p*, u*, rho*L, rho*R (n x 4)
// Grids are (Ny+2g) x (Nx+2g) cells of 4 values, row-major.
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "gfm2d.hpp"

using namespace gfm2d;

namespace {

struct Reader {
    std::vector<double> v;
    size_t k = 0;
    double next() {
        if (k >= v.size()) throw std::runtime_error("input too short");
        return v[k++];
    }
    int inext() { return (int)next(); }
    void grid(grideuler2type& g) {
        for (auto& row : g)
            for (auto& c : row)
                for (int q = 0; q < 4; ++q) c[q] = next();
    }
};

void put(std::vector<double>& out, const grideuler2type& g) {
    for (const auto& row : g)
        for (const auto& c : row) out.insert(out.end(), c.begin(), c.end());
}

const char* bc_name(int b) {
    return b == 0 ? "transmissive" : (b == 1 ? "reflective" : "periodic");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: shim_check in.bin out.bin\n");
        return 2;
    }
    try {
        Reader r;
        {
            std::ifstream in(argv[1], std::ios::binary | std::ios::ate);
            if (!in) throw std::runtime_error("cannot open input");
            const std::streamsize n = in.tellg();
            in.seekg(0);
            r.v.resize((size_t)n / sizeof(double));
            in.read(reinterpret_cast<char*>(r.v.data()), n);
        }
        const int mode = r.inext();
        sim_info p;
        p.Nx = r.inext(); p.Ny = r.inext(); p.numGC = r.inext();
        p.x0 = r.next(); p.y0 = r.next(); p.dx = r.next(); p.dy = r.next();
        p.BC_L = bc_name(r.inext()); p.BC_R = bc_name(r.inext());
        p.BC_B = bc_name(r.inext()); p.BC_T = bc_name(r.inext());
        const int H = p.Ny + 2 * p.numGC, W = p.Nx + 2 * p.numGC;
        std::vector<double> out;
        if (mode == 1) {
            const int scheme = r.inext();
            const double gamma = r.next(), pinf = r.next(), dt = r.next();
            const int b0 = r.inext(), b1 = r.inext(), b2 = r.inext(), b3 = r.inext();
            grideuler2type grid = make_grid(p), future = make_grid(p);
            r.grid(grid);
            std::shared_ptr<pure_RS_base> rs = std::make_shared<pure_RS_exact>();
            std::unique_ptr<FS_base> fs;
            if (scheme == 0) fs.reset(new FS_MUSCL(rs));
            else fs.reset(new FS_godunov(rs));
            fs->pure_fluid_update(p, gamma, pinf, dt, BBrange(b0, b1, b2, b3), grid, future);
            put(out, grid);
            put(out, future);
        } else if (mode == 2) {
            const int variant = r.inext(), ext = r.inext(), slip = r.inext();
            binarySGparams e;
            e.gamma1 = r.next(); e.pinf1 = r.next(); e.gamma2 = r.next(); e.pinf2 = r.next();
            const double t = r.next();
            Levelset ls(p, 5);
            for (auto& x : ls.phi) x = r.next();
            grideuler2type g1 = make_grid(p), g2 = make_grid(p);
            r.grid(g1);
            r.grid(g2);
            std::unique_ptr<GFM_base> gfm;
            switch (variant) {
                case 0: gfm.reset(new GFM_original(p)); break;
                case 1: gfm.reset(new GFM_real(p)); break;
                case 2: gfm.reset(new GFM_riemann(p)); break;
                default: gfm.reset(new GFM_modified(p)); break;
            }
            gfm->use_extension_vfield = ext != 0;
            gfm->use_slip_BCs = slip != 0;
            BBrange rc1(0, 0, 0, 0), rc2(0, 0, 0, 0);
            gfm->set_ghost_states(p, ls, e, t, g1, g2, rc1, rc2);
            put(out, g1);
            put(out, g2);
            for (const BBrange* b : {&rc1, &rc2}) {
                out.push_back(b->imin); out.push_back(b->imax); out.push_back(b->jmin); out.push_back(b->jmax);
            }
            for (int i = 0; i < H; ++i)
                for (int j = 0; j < W; ++j) {
                    out.push_back(gfm->vfield->v[i][j][0]);
                    out.push_back(gfm->vfield->v[i][j][1]);
                }
        } else if (mode == 3) {
            const int N = r.inext();
            Levelset ls(p, 5);
            for (auto& x : ls.phi) x = r.next();
            grideuler2type g1 = make_grid(p), g2 = make_grid(p);
            r.grid(g1);
            r.grid(g2);
            extrapolate_vector(p, ls, N, g1, g2);
            put(out, g1);
            put(out, g2);
        } else if (mode == 4) {
            grideuler2type g = make_grid(p);
            r.grid(g);
            apply_BCs_euler(p, g);
            put(out, g);
        } else if (mode == 5) {
            const int order = r.inext(), every = r.inext(), iters = r.inext();
            const double t = r.next(), dt = r.next();
            Levelset ls(p, order, every, iters);
            for (auto& x : ls.phi) x = r.next();
            gridVector2dtype v(H, std::vector<vec2>(W));
            for (int i = 0; i < H; ++i)
                for (int j = 0; j < W; ++j) {
                    v[i][j][0] = r.next();
                    v[i][j][1] = r.next();
                }
            solved_velocity_field vf;
            vf.store_velocity_field(v, t);
            ls.update_interface_coupled(vf, t, dt);
            out = ls.phi;
        } else if (mode == 6 || mode == 7) {
            const bool pure = mode == 6;
            const int px = pure ? r.inext() : 0;
            const double a = r.next(), b = r.next();
            const double c = pure ? 0.0 : r.next(), d = pure ? 0.0 : r.next();
            const int n = r.inext();
            std::vector<vec4type> L(n), R(n);
            for (auto& s : L)
                for (auto& x : s) x = r.next();
            for (auto& s : R)
                for (auto& x : s) x = r.next();
            pure_RS_exact prs;
            mixed_RS_exact mrs;
            for (int k = 0; k < n; ++k) {
                if (pure) {
                    vec4type f;
                    prs.solve_RP(px != 0, a, b, L[k], R[k], f);
                    out.insert(out.end(), f.begin(), f.end());
                } else {
                    double ps, us, rl, rr;
                    mrs.solve_mixed_RP(a, b, c, d, L[k], R[k], ps, us, rl, rr);
                    out.insert(out.end(), {ps, us, rl, rr});
                }
            }
        } else {
            throw std::runtime_error("unknown mode");
        }
        std::ofstream o(argv[2], std::ios::binary);
        o.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
        return o ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "shim_check: %s\n", e.what());
        return 1;
    }
}
