// gfm2d.cpp -- implementation of the reference-shaped C++ host layer over the
// C ABI.  Every numerical phase runs on the GPU; host code only converts
// between the reference's nested grids and the ABI's flat arrays.
#include "gfm2d.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace gfm2d {

void check(int rc, const char* what) {
    if (rc != GFM_OK) throw std::runtime_error(std::string(what) + ": " + gfm_last_error());
}

static int bc_code(const std::string& s) {
    if (s == "transmissive") return GFM_BC_TRANSMISSIVE;
    if (s == "reflective") return GFM_BC_REFLECTIVE;
    if (s == "periodic") return GFM_BC_PERIODIC;
    throw std::runtime_error("[boundary_conditions] Invalid BC " + s);
}

gfm_sim_info sim_info::c_info() const {
    gfm_sim_info s;
    s.Nx = Nx; s.Ny = Ny; s.numGC = numGC;
    s.x0 = x0; s.y0 = y0; s.dx = dx; s.dy = dy;
    s.BC_L = bc_code(BC_L); s.BC_R = bc_code(BC_R); s.BC_B = bc_code(BC_B); s.BC_T = bc_code(BC_T);
    return s;
}

// GFM_settingsfile.hpp:24-59
void GFM_settingsfile::read_settings_file(const std::string& filename) {
    std::ifstream infile(filename);
    if (!infile) throw std::runtime_error("cannot open settings file " + filename);
    std::string line;
    while (std::getline(infile, line)) {
        std::istringstream iss(line);
        std::string key;
        iss >> key;
        if (key == "Nx") iss >> Nx;
        else if (key == "Ny") iss >> Ny;
        else if (key == "CFL") iss >> CFL;
        else if (key == "sim_type") iss >> sim_type;
        else if (key == "FS") iss >> FS;
        else if (key == "test_case") iss >> test_case;
        else if (key == "ITM") iss >> ITM;
        else if (key == "GFM") iss >> GFM;
        else if (key == "outputpath") iss >> outputpath;
        else if (key == "steps") iss >> steps;
        else if (key == "precision") iss >> precision;
        else if (key == "reinit_every") iss >> reinit_every;
        else if (key == "reinit_iters") iss >> reinit_iters;
        else if (key == "bubble_gamma") iss >> bubble_gamma;
        else if (key == "bubble_rho") iss >> bubble_rho;
        else if (key == "mach") iss >> mach;
        else if (key == "radius") iss >> radius;
        else if (key == "amp") iss >> amp;
        else if (key == "modes") iss >> modes;
    }
    if (sim_type == "onefluid")
        basename = outputpath + sim_type + "-" + test_case + "-" + FS + "-" + std::to_string(Nx) + "-" + std::to_string(Ny);
    else
        basename = outputpath + GFM + "-" + ITM + "-" + test_case + "-" + FS + "-" + std::to_string(Nx) + "-" +
                   std::to_string(Ny);
}

grideuler2type make_grid(const sim_info& p) {
    return grideuler2type(p.Ny + 2 * p.numGC, roweuler2type(p.Nx + 2 * p.numGC, vec4type{0.0, 0.0, 0.0, 0.0}));
}

// ---------------------------------------------------------------- flat <-> nested
static std::vector<double> flatten(const grideuler2type& g) {
    std::vector<double> out;
    out.reserve(g.size() * (g.empty() ? 0 : g[0].size()) * 4);
    for (const auto& row : g)
        for (const auto& c : row) out.insert(out.end(), c.begin(), c.end());
    return out;
}

static void unflatten(const std::vector<double>& f, grideuler2type& g) {
    size_t q = 0;
    for (auto& row : g)
        for (auto& c : row)
            for (int k = 0; k < 4; ++k) c[k] = f[q++];
}

static std::vector<double> ls_flat(const sim_info& p, const GFM_ITM_interface& ls) {
    if (auto lv = dynamic_cast<const Levelset*>(&ls)) return lv->phi;
    const int H = p.Ny + 2 * p.numGC, W = p.Nx + 2 * p.numGC;
    std::vector<double> phi((size_t)H * W);
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) phi[(size_t)i * W + j] = ls.get_sdf(i, j);
    return phi;
}

// ---------------------------------------------------------------- context cache
// Phase-by-phase drop-ins reuse one committed context per (methods, grid, EOS).
// ext, slip: GFM_base::use_extension_vfield / use_slip_BCs (two-fluid contexts)
static gfm_ctx* cached_ctx(const gfm_methods& m, const sim_info& p, const gfm_eos& e, bool ext = false,
                           bool slip = false) {
    static std::map<std::string, gfm_ctx*> cache;   // contexts live until exit
    char key[512];
    const gfm_sim_info s = p.c_info();
    std::snprintf(key, sizeof key, "%d %d %d %d %d %d %d|%d %d %d %.17g %.17g %.17g %.17g %d %d %d %d|%.17g %.17g %.17g %.17g|%d %d",
                  m.sim_type, m.fs, m.gfm, m.ls_order, m.reinit_every, m.reinit_iters, m.precision, s.Nx, s.Ny,
                  s.numGC, s.x0, s.y0, s.dx, s.dy, s.BC_L, s.BC_R, s.BC_B, s.BC_T, e.gamma1, e.pinf1, e.gamma2,
                  e.pinf2, ext ? 1 : 0, slip ? 1 : 0);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    gfm_ctx* c = nullptr;
    check(gfm_ctx_create(&m, &c), "gfm_ctx_create");
    if (ext || slip) check(gfm_set_gfm_flags(c, ext ? 1 : 0, slip ? 1 : 0), "gfm_set_gfm_flags");
    check(gfm_instance_add(c, &s, &e, 0.5, 1e300, 1 << 30, nullptr), "gfm_instance_add");
    check(gfm_ctx_commit(c), "gfm_ctx_commit");
    cache[key] = c;
    return c;
}

static gfm_methods methods_of(int sim_type, int fs, int gfm, int order = 5, int every = 0, int iters = 0,
                              int precision = GFM_FP64, int itm = GFM_ITM_LS) {
    gfm_methods m;
    m.sim_type = sim_type; m.fs = fs; m.gfm = gfm; m.ls_order = order;
    m.reinit_every = every; m.reinit_iters = iters; m.precision = precision; m.itm = itm;
    return m;
}

// ---------------------------------------------------------------- Riemann solvers
void pure_RS_exact::solve_RP(bool paralleltox, double gamma, double pinf, const vec4type& L, const vec4type& R,
                             vec4type& flux) const {
    unsigned st = 0;
    check(gfm_pure_rs_solve(1, paralleltox ? 1 : 0, gamma, pinf, L.data(), R.data(), flux.data(), &st),
          "gfm_pure_rs_solve");
}

void mixed_RS_exact::solve_mixed_RP(double gL, double pL, double gR, double pR, const vec4type& L, const vec4type& R,
                                    double& p_star, double& u_star, double& rho_star_L, double& rho_star_R) const {
    double out[4];
    unsigned st = 0;
    check(gfm_mixed_rs_solve(1, gL, pL, gR, pR, L.data(), R.data(), out, &st), "gfm_mixed_rs_solve");
    p_star = out[0];
    u_star = out[1];
    rho_star_L = out[2];
    rho_star_R = out[3];
    if (p_star != p_star) std::printf("Exact RS returned NaN\n");
}

// ---------------------------------------------------------------- flux solvers
static void fs_update(int scheme, const sim_info& params, double gamma, double pinf, double dt,
                      const BBrange& rc, grideuler2type& grid, grideuler2type& future) {
    gfm_eos e{gamma, pinf, gamma, pinf};
    gfm_ctx* c = cached_ctx(methods_of(GFM_SIM_ONEFLUID, scheme, GFM_RIEMANN), params, e);
    std::vector<double> g = flatten(grid);
    check(gfm_upload_state(c, 0, g.data(), nullptr, nullptr), "gfm_upload_state");
    check(gfm_set_dt(c, 0, dt), "gfm_set_dt");
    const int bb[4] = {rc.imin, rc.imax, rc.jmin, rc.jmax};
    check(gfm_set_bbox(c, 0, bb, bb), "gfm_set_bbox");
    check(gfm_fs_update(c), "gfm_fs_update");
    check(gfm_download_state(c, 0, g.data(), nullptr, nullptr), "gfm_download_state");
    unflatten(g, grid);
    for (int i = rc.imin; i < rc.imax; ++i)
        for (int j = rc.jmin; j < rc.jmax; ++j) future[i][j] = grid[i][j];
}

void FS_MUSCL::pure_fluid_update(const sim_info& p, const double gamma, const double pinf, const double dt,
                                 const BBrange& rc, grideuler2type& grid, grideuler2type& future) {
    fs_update(GFM_FS_MUSCL, p, gamma, pinf, dt, rc, grid, future);
}

void FS_godunov::pure_fluid_update(const sim_info& p, const double gamma, const double pinf, const double dt,
                                   const BBrange& rc, grideuler2type& grid, grideuler2type& future) {
    fs_update(GFM_FS_GODUNOV, p, gamma, pinf, dt, rc, grid, future);
}

// ---------------------------------------------------------------- level set
Levelset::Levelset(const sim_info& p, int order, int every, int iters)
    : phi((size_t)(p.Ny + 2 * p.numGC) * (p.Nx + 2 * p.numGC), 0.0), params(p), order(order),
      reinit_every(every), reinit_iters(iters), W(p.Nx + 2 * p.numGC) {}

vec2 Levelset::get_normal(int i, int j) const {
    return {(get_sdf(i, j + 1) - get_sdf(i, j - 1)) / (2.0 * params.dx),
            (get_sdf(i + 1, j) - get_sdf(i - 1, j)) / (2.0 * params.dy)};
}

vec2 Levelset::get_normal(const vec2& pos) const {
    const int j = (int)std::floor((pos[0] - params.x0) / params.dx) + params.numGC;
    const int i = (int)std::floor((pos[1] - params.y0) / params.dy) + params.numGC;
    const int H = params.Ny + 2 * params.numGC;
    if (i < 1 || j < 1 || i > H - 2 || j > W - 2) throw std::runtime_error("get_normal: position outside the grid");
    return get_normal(i, j);
}

bool Levelset::is_interfacial_cell(int i, int j) const {
    const bool s = get_sdf(i, j) <= 0.0;
    return ((get_sdf(i, j + 1) <= 0.0) != s) || ((get_sdf(i, j - 1) <= 0.0) != s) ||
           ((get_sdf(i + 1, j) <= 0.0) != s) || ((get_sdf(i - 1, j) <= 0.0) != s);
}

void Levelset::update_interface_coupled(const solved_velocity_field& vf, double t, double dt) {
    gfm_eos e{1.4, 0.0, 1.4, 0.0};
    gfm_ctx* c = cached_ctx(methods_of(GFM_SIM_TWOFLUID, GFM_FS_MUSCL, GFM_RIEMANN, order, reinit_every,
                                       reinit_iters), params, e);
    const size_t n = phi.size();
    std::vector<double> dummy(n * 4), v(n * 2);
    for (size_t q = 0; q < n; ++q) { dummy[4 * q] = 1.0; dummy[4 * q + 3] = 2.5; }
    for (size_t i = 0; i < vf.v.size(); ++i)
        for (size_t j = 0; j < vf.v[i].size(); ++j) {
            v[2 * (i * W + j)] = vf.v[i][j][0];
            v[2 * (i * W + j) + 1] = vf.v[i][j][1];
        }
    check(gfm_upload_state(c, 0, dummy.data(), dummy.data(), phi.data()), "gfm_upload_state");
    check(gfm_upload_vfield(c, 0, v.data()), "gfm_upload_vfield");
    check(gfm_set_time(c, 0, t, nsteps), "gfm_set_time");
    check(gfm_set_dt(c, 0, dt), "gfm_set_dt");
    check(gfm_ls_update(c), "gfm_ls_update");
    check(gfm_download_state(c, 0, nullptr, nullptr, phi.data()), "gfm_download_state");
    nsteps++;
}

// ---------------------------------------------------------------- ghost fluid methods
void GFM_base::run_variant(int variant, const sim_info& p, const GFM_ITM_interface& ls, const binarySGparams& eos,
                           double t, grideuler2type& grid1, grideuler2type& grid2, BBrange& rc1, BBrange& rc2) {
    gfm_eos e{eos.gamma1, eos.pinf1, eos.gamma2, eos.pinf2};
    gfm_ctx* c = cached_ctx(methods_of(GFM_SIM_TWOFLUID, GFM_FS_MUSCL, variant), p, e, use_extension_vfield,
                            use_slip_BCs);
    std::vector<double> g1 = flatten(grid1), g2 = flatten(grid2), phi = ls_flat(p, ls);
    check(gfm_upload_state(c, 0, g1.data(), g2.data(), phi.data()), "gfm_upload_state");
    check(gfm_set_ghost_states(c), "gfm_set_ghost_states");
    check(gfm_download_state(c, 0, g1.data(), g2.data(), nullptr), "gfm_download_state");
    unflatten(g1, grid1);
    unflatten(g2, grid2);
    std::vector<double> v(phi.size() * 2);
    check(gfm_download_vfield(c, 0, v.data()), "gfm_download_vfield");
    const int H = p.Ny + 2 * p.numGC, W = p.Nx + 2 * p.numGC;
    gridVector2dtype nv(H, std::vector<vec2>(W));
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) nv[i][j] = {v[2 * ((size_t)i * W + j)], v[2 * ((size_t)i * W + j) + 1]};
    vfield->store_velocity_field(nv, t);
    int b1[4], b2[4];
    check(gfm_get_bbox(c, 0, b1, b2), "gfm_get_bbox");
    rc1 = BBrange(b1[0], b1[1], b1[2], b1[3]);
    rc2 = BBrange(b2[0], b2[1], b2[2], b2[3]);
}

void extrapolate_vector(const sim_info& p, const GFM_ITM_interface& ls, const int N, grideuler2type& grid1,
                        grideuler2type& grid2) {
    gfm_eos e{1.4, 0.0, 1.4, 0.0};
    gfm_ctx* c = cached_ctx(methods_of(GFM_SIM_TWOFLUID, GFM_FS_MUSCL, GFM_RIEMANN), p, e);
    std::vector<double> g1 = flatten(grid1), g2 = flatten(grid2), phi = ls_flat(p, ls);
    check(gfm_upload_state(c, 0, g1.data(), g2.data(), phi.data()), "gfm_upload_state");
    check(gfm_extrapolate(c, N), "gfm_extrapolate");
    check(gfm_download_state(c, 0, g1.data(), g2.data(), nullptr), "gfm_download_state");
    unflatten(g1, grid1);
    unflatten(g2, grid2);
}

void apply_BCs_euler(const sim_info& p, grideuler2type& grid) {
    gfm_eos e{1.4, 0.0, 1.4, 0.0};
    gfm_ctx* c = cached_ctx(methods_of(GFM_SIM_ONEFLUID, GFM_FS_MUSCL, GFM_RIEMANN), p, e);
    std::vector<double> g = flatten(grid);
    check(gfm_upload_state(c, 0, g.data(), nullptr, nullptr), "gfm_upload_state");
    check(gfm_apply_bcs(c), "gfm_apply_bcs");
    check(gfm_download_state(c, 0, g.data(), nullptr, nullptr), "gfm_download_state");
    unflatten(g, grid);
}

// ---------------------------------------------------------------- simulations
// sim_onefluid::set_sim_parameters (sim_onefluid.cpp:96-146)
void set_onefluid_parameters(const GFM_settingsfile& SF, sim_info& p, binarySGparams& e) {
    p.numGC = 2; p.Nx = SF.Nx; p.Ny = SF.Ny; p.output_freq = 1000000000;
    e.gamma1 = 1.4; e.gamma2 = 0.0; e.pinf1 = 0.0; e.pinf2 = 0.0;
    p.x0 = 0.0; p.y0 = 0.0; p.dx = 1.0 / p.Nx; p.dy = 1.0 / p.Ny;
    p.BC_L = p.BC_R = p.BC_B = p.BC_T = "transmissive";
    const std::string& tc = SF.test_case;
    if (tc == "TTC1_x" || tc == "TTC1_y") p.T = 0.25;
    else if (tc == "TTC2_x" || tc == "TTC2_y") p.T = 0.15;
    else if (tc == "TTC3_x" || tc == "TTC3_y") p.T = 0.012;
    else if (tc == "TTC4_x" || tc == "TTC4_y" || tc == "TTC5_x" || tc == "TTC5_y") p.T = 0.035;
    else if (tc == "circularexplosion") { p.dx = 2.0 / p.Nx; p.dy = 2.0 / p.Ny; p.T = 0.25; }
    else throw std::runtime_error("Invalid test_case choice.");
}

static int onefluid_case_id(const std::string& tc) {
    if (tc == "circularexplosion") return 100;
    const int k = tc[3] - '0';
    return tc.back() == 'y' ? 10 + k : k;
}

static void write_prims(const std::string& fname, const sim_info& p, const binarySGparams& e,
                        const std::vector<double>& g1, const std::vector<double>* g2, const std::vector<double>* phi,
                        int i0, int i1, int j0, int j1) {
    std::ofstream out(fname);
    const int W = p.Nx + 2 * p.numGC;
    for (int i = i0; i < i1; ++i) {
        for (int j = j0; j < j1; ++j) {
            const vec2 cc = p.cellcentre_coord(i, j);
            const size_t q = (size_t)i * W + j;
            const bool one = !phi || (*phi)[q] < 0.0;
            const double* U = one ? &g1[4 * q] : &(*g2)[4 * q];
            const double g = one ? e.gamma1 : e.gamma2, pinf = one ? e.pinf1 : e.pinf2;
            const double rho = U[0], u = U[1] / rho, v = U[2] / rho;
            const double en = U[3] / rho - 0.5 * (u * u + v * v);
            const double pr = (g - 1.0) * rho * en - g * pinf;
            out << cc[0] << " " << cc[1] << " " << rho << " " << u << " " << v << " " << en << " " << pr << "\n";
        }
        out << "\n";
    }
}

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// loop body of sim_onefluid::run_sim (sim_onefluid.cpp:12-64) on the device
void sim_onefluid::run_sim(GFM_settingsfile SF) {
    sim_info p;
    binarySGparams e;
    set_onefluid_parameters(SF, p, e);
    const int fs = SF.FS == "Godunov" ? GFM_FS_GODUNOV : (SF.FS == "MUSCL" ? GFM_FS_MUSCL : -1);
    if (fs < 0) throw std::runtime_error("Invalid method choice");
    gfm_methods m = methods_of(GFM_SIM_ONEFLUID, fs, GFM_RIEMANN, 5, 0, 0,
                               SF.precision == "mixed" ? GFM_MIXED : GFM_FP64);
    gfm_ctx* c = nullptr;
    check(gfm_ctx_create(&m, &c), "gfm_ctx_create");
    const gfm_sim_info s = p.c_info();
    const gfm_eos ce{e.gamma1, e.pinf1, e.gamma1, e.pinf1};
    check(gfm_instance_add(c, &s, &ce, SF.CFL, p.T, SF.steps > 0 ? SF.steps : 1 << 30, nullptr), "gfm_instance_add");
    check(gfm_ctx_commit(c), "gfm_ctx_commit");
    const size_t n = (size_t)(p.Ny + 2 * p.numGC) * (p.Nx + 2 * p.numGC);
    std::vector<double> g(4 * n);
    check(gfm_ic_onefluid(&s, &ce, onefluid_case_id(SF.test_case), g.data()), "gfm_ic_onefluid");
    check(gfm_upload_state(c, 0, g.data(), nullptr, nullptr), "gfm_upload_state");
    // output at step 0 and at the end (:34, :60); the per-step test
    // numsteps % output_freq (:54) never holds with output_freq 10^9
    write_prims(SF.basename + "-0-prims.dat", p, e, g, nullptr, nullptr, 0, p.Ny + 2 * p.numGC, 0,
                p.Nx + 2 * p.numGC);
    std::printf("[%s] Initialisation complete. Beginning time iterations with CFL = %g.\n", SF.basename.c_str(), SF.CFL);
    const double t0 = now();
    double t = 0.0, dt = 0.0;
    int nstep = 0, prev = -1;
    while (nstep != prev) {
        prev = nstep;
        check(gfm_step_graph(c, 50), "gfm_step_graph");
        check(gfm_get_time(c, 0, &t, &nstep, &dt), "gfm_get_time");
    }
    const double t1 = now();
    unsigned st = 0;
    check(gfm_status(c, 0, &st), "gfm_status");
    check(gfm_download_state(c, 0, g.data(), nullptr, nullptr), "gfm_download_state");
    write_prims(SF.basename + "-" + std::to_string(nstep) + "-prims.dat", p, e, g, nullptr, nullptr, 0,
                p.Ny + 2 * p.numGC, 0, p.Nx + 2 * p.numGC);
    std::printf("[%s] Simulation complete: %d steps, t = %.10g, status %u.\n", SF.basename.c_str(), nstep, t, st);
    std::printf("[%s] Simulation took %gs.\n", SF.basename.c_str(), t1 - t0);
    check(gfm_ctx_destroy(c), "gfm_ctx_destroy");
}

// sim_twofluid::set_sim_parameters (sim_twofluid.cpp:336-490) for the level-set
// test cases of the hot path, plus the SI shocked bubble and the periodic RMI
void set_twofluid_parameters(const GFM_settingsfile& SF, sim_info& p, binarySGparams& e) {
    p.numGC = 8; p.Nx = SF.Nx; p.Ny = SF.Ny; p.output_freq = 4;
    e = binarySGparams();
    p.x0 = 0.0; p.y0 = 0.0; p.dx = 1.0 / p.Nx; p.dy = 1.0 / p.Ny;
    p.BC_L = p.BC_R = p.BC_B = p.BC_T = "transmissive";
    const std::string& tc = SF.test_case;
    if (tc == "circularexplosion") { p.dx = 2.0 / p.Nx; p.dy = 2.0 / p.Ny; p.T = 0.25; }
    else if (tc == "ST1_x") { e.gamma2 = 1.2; p.T = 0.0007; }
    else if (tc == "ST2_y") {
        e.gamma1 = 4.4; e.pinf1 = 6e8; p.x0 = -2.0; p.y0 = -2.0; p.dx = 4.0 / p.Nx; p.dy = 4.0 / p.Ny; p.T = 0.0009;
    } else if (tc == "shocked_helium_bubble") {
        e.gamma1 = 1.667; p.dx = 325.0 / p.Nx; p.dy = 89.0 / p.Ny; p.T = 280.0;
        p.BC_T = p.BC_B = "reflective"; p.output_freq = 2;
    } else if (tc == "RMI" || tc == "RMI_pert" || tc == "RMI_unpert") {   // :430-442
        e.gamma2 = 1.093; p.dx = 4.0 / p.Nx; p.dy = 1.0 / p.Ny; p.T = tc == "RMI_pert" ? 4.0 : 10.0;
        p.BC_T = p.BC_B = "reflective"; p.output_freq = 10;
    } else if (tc == "shocked_bubble") {          // SI units, parameterised gas
        e.gamma1 = SF.bubble_gamma; p.dx = 0.325 / p.Nx; p.dy = 0.089 / p.Ny; p.T = 1e300;
        p.BC_T = p.BC_B = "reflective";
    } else if (tc == "RMI_periodic") {            // single mode, periodic in y
        e.gamma2 = 1.093; p.dx = 4.0 / p.Nx; p.dy = 1.0 / p.Ny; p.T = 1e300;
        p.BC_T = p.BC_B = "periodic";
    } else {
        throw std::runtime_error("Invalid test_case choice.");
    }
}

// sim_twofluid::output (sim_twofluid.cpp:163-295): <stem>-prims.dat,
// -ls.dat, -vfield.dat, -z.dat, -schlieren.dat and <base>-GDAerr.dat
static void write_twofluid_output(const std::string& stem, const std::string& base, const sim_info& p,
                                  const binarySGparams& e, const std::vector<double>& g1,
                                  const std::vector<double>& g2, const std::vector<double>& phi,
                                  const std::vector<double>& vf, const std::vector<double>& z) {
    const int H = p.Ny + 2 * p.numGC, W = p.Nx + 2 * p.numGC, g = p.numGC;
    write_prims(stem + "-prims.dat", p, e, g1, &g2, &phi, 2, H - 2, 2, W - 2);
    {
        std::ofstream ls(stem + "-ls.dat"), v(stem + "-vfield.dat"), zo(stem + "-z.dat");
        for (int i = 2; i < H - 2; ++i) {
            for (int j = 2; j < W - 2; ++j) {
                const vec2 cc = p.cellcentre_coord(i, j);
                const size_t q = (size_t)i * W + j;
                ls << cc[0] << " " << cc[1] << " " << phi[q] << "\n";
                v << cc[0] << " " << cc[1] << " " << vf[2 * q] << " " << vf[2 * q + 1] << "\n";
                zo << cc[0] << " " << cc[1] << " " << z[q] << "\n";
            }
            ls << "\n";
            v << "\n";
            zo << "\n";
        }
    }
    // numerical schlieren (:226-271): |grad rho| by central differences over
    // the interior, each neighbour's density from the fluid real there
    // (phi <= 0: fluid 1), over its maximum
    auto rho = [&](int i, int j) {
        const size_t q = (size_t)i * W + j;
        return phi[q] <= 0.0 ? g1[4 * q] : g2[4 * q];
    };
    std::vector<double> grad;
    grad.reserve((size_t)p.Nx * p.Ny);
    for (int i = g; i < p.Ny + g; ++i)
        for (int j = g; j < p.Nx + g; ++j) {
            const double gx = (rho(i, j + 1) - rho(i, j - 1)) / (2.0 * p.dx);
            const double gy = (rho(i + 1, j) - rho(i - 1, j)) / (2.0 * p.dy);
            grad.push_back(std::sqrt(gx * gx + gy * gy));
        }
    const double gmax = *std::max_element(grad.begin(), grad.end());
    {
        std::ofstream out(stem + "-schlieren.dat");
        size_t k = 0;
        for (int i = g; i < p.Ny + g; ++i) {
            for (int j = g; j < p.Nx + g; ++j) {
                const vec2 cc = p.cellcentre_coord(i, j);
                out << cc[0] << " " << cc[1] << " " << grad[k++] / gmax << "\n";
            }
            out << "\n";
        }
    }
    // L1 distance of fluid 1's density to the Gaussian density bump of the
    // advection test, 1 + 1000 exp(-|x - (0.5, 0.5)|^2 / (2 0.1^2)), written
    // for every case and overwritten at every output, as the reference does
    // (:273-291)
    double l1 = 0.0;
    for (int i = g; i < p.Ny + g; ++i)
        for (int j = g; j < p.Nx + g; ++j) {
            const vec2 cc = p.cellcentre_coord(i, j);
            const double r2 = (cc[0] - 0.5) * (cc[0] - 0.5) + (cc[1] - 0.5) * (cc[1] - 0.5);
            const double exact = 1.0 + 1000.0 * std::exp(-r2 / (2.0 * 0.1 * 0.1));
            l1 += std::fabs(g1[4 * ((size_t)i * W + j)] - exact) * p.dx * p.dy;
        }
    std::ofstream err(base + "-GDAerr.dat");
    err << p.Nx << " " << l1 << "\n";
}

// the level set's HOUC order for settings ITM (sim_twofluid.cpp:555-589); CLSVOF carries a fifth-order
// level set beside its volume fraction
static int ls_order(const std::string& itm) {
    if (itm == "LS1") return 1;
    if (itm == "LS3") return 3;
    if (itm == "LS5" || itm == "CLSVOF") return 5;
    if (itm == "LS9") return 9;
    throw std::runtime_error("Invalid ITM choice (LS1/LS3/LS5/LS9, CLSVOF; the PYVOF, MOF and EMOF2 trackers are "
                             "not built)");
}

static int gfm_id(const std::string& g) {
    if (g == "OGFM") return GFM_ORIGINAL;
    if (g == "realGFM") return GFM_REAL;
    if (g == "riemannGFM") return GFM_RIEMANN;
    if (g == "MGFM") return GFM_MODIFIED;
    if (g == "VOFGFM") return GFM_VOFGFM;
    if (g == "OGFMVOF") return GFM_OGFMVOF;
    throw std::runtime_error("Invalid GFM choice (OGFM, realGFM, riemannGFM, MGFM, VOFGFM, OGFMVOF)");
}

// sim_twofluid::run_sim (sim_twofluid.cpp:27-103) with the whole loop on the device
void sim_twofluid::run_sim(GFM_settingsfile SF) {
    sim_info p;
    binarySGparams e;
    set_twofluid_parameters(SF, p, e);
    const int fs = SF.FS == "Godunov" ? GFM_FS_GODUNOV : (SF.FS == "MUSCL" ? GFM_FS_MUSCL : -1);
    if (fs < 0) throw std::runtime_error("Invalid FS choice");
    gfm_methods m = methods_of(GFM_SIM_TWOFLUID, fs, gfm_id(SF.GFM), ls_order(SF.ITM), SF.reinit_every,
                               SF.reinit_iters, SF.precision == "mixed" ? GFM_MIXED : GFM_FP64,
                               SF.ITM == "CLSVOF" ? GFM_ITM_CLSVOF : GFM_ITM_LS);
    gfm_ctx* c = nullptr;
    check(gfm_ctx_create(&m, &c), "gfm_ctx_create");
    const gfm_sim_info s = p.c_info();
    const gfm_eos ce{e.gamma1, e.pinf1, e.gamma2, e.pinf2};
    const int max_steps = SF.steps > 0 ? SF.steps : (1 << 30);
    check(gfm_instance_add(c, &s, &ce, SF.CFL, p.T, max_steps, nullptr), "gfm_instance_add");
    // the two points of output_perturbation_amplitude (:1852-1853), kept on
    // their rows (:1882-1884)
    const bool tracers = SF.test_case == "RMI_pert";
    const double xy[4] = {2.88, 1.0, 2.92, 0.5};
    if (tracers) check(gfm_set_tracers(c, 0, 2, xy, 1), "gfm_set_tracers");
    check(gfm_ctx_commit(c), "gfm_ctx_commit");
    const size_t n = (size_t)(p.Ny + 2 * p.numGC) * (p.Nx + 2 * p.numGC);
    std::vector<double> g1(4 * n), g2(4 * n), phi(n);
    const std::string& tc = SF.test_case;
    int init_reinit = 0;
    if (tc == "circularexplosion") {          // sim_twofluid.cpp:609-673
        const double L[4] = {1.0, 0.0, 0.0, 1.0}, R[4] = {0.125, 0.0, 0.0, 0.1};
        check(gfm_ic_bubble(&s, &ce, 1.0, 1.0, 0.4, L, R, R, 1e300, g1.data(), g2.data(), phi.data()), "ic");
    } else if (tc == "ST1_x") {
        const double L[4] = {1.0, 0.0, 0.0, 1e5}, R[4] = {0.125, 0.0, 0.0, 1e4};
        check(gfm_ic_planar(&s, &ce, 0, 0.5, L, R, g1.data(), g2.data(), phi.data()), "ic");
    } else if (tc == "ST2_y") {
        const double L[4] = {1000.0, 0.0, 0.0, 1e9}, R[4] = {50.0, 0.0, 0.0, 1e5};
        check(gfm_ic_planar(&s, &ce, 1, 0.7, L, R, g1.data(), g2.data(), phi.data()), "ic");
    } else if (tc == "shocked_helium_bubble") {   // sim_twofluid.cpp:774-859
        const double he[4] = {0.138, 0.0, 0.0, 1.0}, pre[4] = {1.0, 0.0, 0.0, 1.0};
        const double post[4] = {1.3764, -0.394, 0.0, 1.5698};
        check(gfm_ic_bubble(&s, &ce, 175.0, 44.5, 25.0, he, pre, post, 225.0, g1.data(), g2.data(), phi.data()), "ic");
    } else if (tc == "RMI" || tc == "RMI_pert" || tc == "RMI_unpert" || tc == "RMI_periodic") {
        // sim_twofluid.cpp:1245-1347 (RMI), :1348-1453 (RMI_pert), :1454-1553 (RMI_unpert): the
        // interface x = 2.9 - eps cos(2 pi y) with eps 0.2, 0.02, 0
        const double air[4] = {1.0, 0.0, 0.0, 1.0}, sf6[4] = {5.04, 0.0, 0.0, 1.0};
        double post[4] = {1.411, -0.39, 0.0, 1.628};
        double amp = tc == "RMI_pert" ? 0.02 : (tc == "RMI_unpert" ? 0.0 : 0.2);
        int modes = 1;
        if (tc == "RMI_periodic") {
            check(gfm_shock_state(1.4, 0.0, air, SF.mach, -1, post), "shock");
            modes = SF.modes;
            amp = SF.amp / modes;
        }
        check(gfm_ic_rmi(&s, &ce, 2.9, amp, modes, air, post, sf6, 3.2, g1.data(), g2.data(), phi.data()), "ic");
        init_reinit = 40;
    } else if (tc == "shocked_bubble") {
        const double pre[4] = {1.0, 0.0, 0.0, 1e5};
        const double bub[4] = {SF.bubble_rho, 0.0, 0.0, 1e5};
        double post[4];
        check(gfm_shock_state(1.4, 0.0, pre, SF.mach, -1, post), "shock");
        check(gfm_ic_bubble(&s, &ce, 0.175, 0.0445, SF.radius, bub, pre, post, 0.225, g1.data(), g2.data(),
                            phi.data()), "ic");
    }
    check(gfm_upload_state(c, 0, g1.data(), g2.data(), phi.data()), "gfm_upload_state");
    check(gfm_ls_reinit(c, init_reinit), "gfm_ls_reinit");   // ls->initialisation() (:47)
    const int H = p.Ny + 2 * p.numGC, W = p.Nx + 2 * p.numGC;
    std::vector<double> vf(2 * (size_t)H * W), z((size_t)H * W);
    // sim_twofluid::output (:163-295) of the current state
    auto output = [&](int nstep) {
        check(gfm_download_state(c, 0, g1.data(), g2.data(), phi.data()), "gfm_download_state");
        // the velocity field of the last ghost fill, the one the level set was
        // advected with (:181,198); zero before the first step
        check(gfm_download_vfield(c, 0, vf.data()), "gfm_download_vfield");
        check(gfm_download_z(c, 0, z.data()), "gfm_download_z");
        write_twofluid_output(SF.basename + "-" + std::to_string(nstep), SF.basename, p, e, g1, g2, phi, vf, z);
    };
    output(0);
    const double interval = p.T / p.output_freq;
    const bool periodic_output = std::isfinite(interval);
    double last = 0.0;
    if (periodic_output) check(gfm_set_output_hold(c, 0, last, interval), "gfm_set_output_hold");
    std::printf("[%s] Initialisation complete. Beginning time iterations with CFL = %g.\n", SF.basename.c_str(), SF.CFL);
    const double t0 = now();
    double t = 0.0, dt = 0.0;
    int nstep = 0, prev = -1;
    // whole graphs of steps; the output hold stops the instance after the step
    // the reference writes its output after (t - lastoutputtime > interval,
    // :79-83), so the host writes it and moves `last` on
    while (nstep != prev) {
        prev = nstep;
        check(gfm_step_graph(c, 50), "gfm_step_graph");
        check(gfm_get_time(c, 0, &t, &nstep, &dt), "gfm_get_time");
        if (periodic_output && t - last > interval) {
            output(nstep);
            last = t;
            check(gfm_set_output_hold(c, 0, last, interval), "gfm_set_output_hold");
            prev = -1;
        }
    }
    const double t1 = now();
    check(gfm_compute_dt(c), "gfm_compute_dt");   // records the final masses
    unsigned st = 0;
    check(gfm_status(c, 0, &st), "gfm_status");
    output(nstep);                                 // :94
    {
        const int nh = std::min({nstep + 1, max_steps + 1, 1 << 20});   // the histories' capacity
        std::vector<double> h(3 * (size_t)nh);
        check(gfm_mass_history(c, 0, nh, h.data()), "gfm_mass_history");
        std::ofstream out(SF.basename + "-masschange.dat");
        for (int k = 0; k < nh; ++k)
            out << h[3 * k] << " " << (h[3 * k + 1] - h[1]) / h[1] << " " << (h[3 * k + 2] - h[2]) / h[2] << "\n";
    }
    if (tracers) {
        // output_perturbation_amplitude (:1831-1887), one line per step:
        // t, the amplitude 0.5 (x1 - x0) and both points' x, then the same
        // for their x velocities
        const int nh = std::min({nstep + 1, max_steps + 1, 1 << 20});   // the histories' capacity
        std::vector<double> h(9 * (size_t)nh);
        check(gfm_tracer_history(c, 0, nh, h.data()), "gfm_tracer_history");
        std::ofstream amp(SF.basename + "-pertamp.dat"), der(SF.basename + "-pertampderiv.dat");
        for (int k = 1; k < nh; ++k) {
            const double* r = &h[9 * (size_t)k];
            amp << r[0] << " " << 0.5 * (r[5] - r[1]) << " " << r[5] << " " << r[1] << "\n";
            der << r[0] << " " << 0.5 * (r[7] - r[3]) << " " << r[7] << " " << r[3] << "\n";
        }
    }
    {
        std::ofstream out(SF.basename + "-runtime.dat");
        out << t1 - t0 << "\n";
    }
    std::printf("[%s] Simulation complete: %d steps, t = %.10g, status %u.\n", SF.basename.c_str(), nstep, t, st);
    std::printf("[%s] Simulation took %gs.\n", SF.basename.c_str(), t1 - t0);
    check(gfm_ctx_destroy(c), "gfm_ctx_destroy");
}

}  // namespace gfm2d
