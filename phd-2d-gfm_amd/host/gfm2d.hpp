// gfm2d.hpp -- C++ host layer with the reference's class interfaces
// (FS_base, pure_RS_base, mixed_RS_base, GFM_base, GFM_ITM_interface,
// sim_base, GFM_settingsfile, extrapolate_vector) over the C ABI of
// libgfm_hip.so.  Each object is a drop-in for the CPU one: it takes the
// reference's host grids, runs the phase on the GPU and writes the result
// back.  sim_onefluid / sim_twofluid keep the whole loop on the device.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "gfm_hip.h"

namespace gfm2d {

// typedefs.hpp of PhD-Common (Eigen vectors there; plain arrays here)
using vec4type = std::array<double, 4>;
using roweuler2type = std::vector<vec4type>;
using grideuler2type = std::vector<roweuler2type>;
using vec2 = std::array<double, 2>;
using gridVector2dtype = std::vector<std::vector<vec2>>;

// sim_info (PhD-Common sim_info.hpp; fields set at sim_twofluid.cpp:343-489)
struct sim_info {
    int Nx = 0, Ny = 0, numGC = 0;
    double x0 = 0.0, y0 = 0.0, dx = 1.0, dy = 1.0, T = 0.0;
    int output_freq = 0;
    std::string BC_L = "transmissive", BC_R = "transmissive", BC_B = "transmissive", BC_T = "transmissive";
    vec2 cellcentre_coord(int i, int j) const { return {x0 + (j - numGC + 0.5) * dx, y0 + (i - numGC + 0.5) * dy}; }
    gfm_sim_info c_info() const;
};

// binarySGparams (sim_twofluid.cpp:347-350)
struct binarySGparams {
    double gamma1 = 1.4, gamma2 = 1.4, pinf1 = 0.0, pinf2 = 0.0;
};

// BBrange.hpp
class BBrange {
public:
    int imin, imax, jmin, jmax;
    BBrange(int imin, int imax, int jmin, int jmax) : imin(imin), imax(imax), jmin(jmin), jmax(jmax) {}
};

// GFM_settingsfile.hpp: the nine keys of the reference plus optional keys for
// the batched / mixed-precision runs (ignored lines stay ignored)
class GFM_settingsfile {
public:
    int Nx = 0, Ny = 0;
    double CFL = 0.0;
    std::string sim_type, FS, ITM, GFM, test_case, outputpath, basename;
    // extensions (not in the reference's file format)
    int steps = 0;                 // stop after this many steps (0: run to T)
    std::string precision = "fp64";
    int reinit_every = 5, reinit_iters = 10;
    double bubble_gamma = 1.648, bubble_rho = 0.182, mach = 1.22, radius = 0.025;
    double amp = 0.05;
    int modes = 1;
    void read_settings_file(const std::string& filename);
};

grideuler2type make_grid(const sim_info& p);

// ---------------------------------------------------------------- Riemann solvers
class pure_RS_base {
public:
    virtual ~pure_RS_base() = default;
    virtual void solve_RP(bool paralleltox, double gamma, double pinf, const vec4type& Lstate,
                          const vec4type& Rstate, vec4type& flux) const = 0;
};

class pure_RS_exact : public pure_RS_base {
public:
    void solve_RP(bool paralleltox, double gamma, double pinf, const vec4type& Lstate, const vec4type& Rstate,
                  vec4type& flux) const override;
};

class mixed_RS_base {
public:
    virtual ~mixed_RS_base() = default;
    virtual void solve_mixed_RP(double gammaL, double pinfL, double gammaR, double pinfR, const vec4type& Lstate,
                                const vec4type& Rstate, double& p_star, double& u_star, double& rho_star_L,
                                double& rho_star_R) const = 0;
};

class mixed_RS_exact : public mixed_RS_base {
public:
    void solve_mixed_RP(double gammaL, double pinfL, double gammaR, double pinfR, const vec4type& Lstate,
                        const vec4type& Rstate, double& p_star, double& u_star, double& rho_star_L,
                        double& rho_star_R) const override;
};

// ---------------------------------------------------------------- flux solvers
class FS_base {
public:
    std::shared_ptr<pure_RS_base> RS;
    explicit FS_base(std::shared_ptr<pure_RS_base> RS) : RS(RS) {}
    virtual ~FS_base() = default;
    virtual void pure_fluid_update(const sim_info& params, const double gamma, const double pinf, const double dt,
                                   const BBrange& realcells, grideuler2type& grid, grideuler2type& future_grid) = 0;
};

class FS_MUSCL : public FS_base {
public:
    using FS_base::FS_base;
    void pure_fluid_update(const sim_info& params, const double gamma, const double pinf, const double dt,
                           const BBrange& realcells, grideuler2type& grid, grideuler2type& future_grid) override;
};

class FS_godunov : public FS_base {
public:
    using FS_base::FS_base;
    void pure_fluid_update(const sim_info& params, const double gamma, const double pinf, const double dt,
                           const BBrange& realcells, grideuler2type& grid, grideuler2type& future_grid) override;
};

// ---------------------------------------------------------------- interface tracking
// solved_velocity_field_firstorder: the velocity field the GFM stores each step
class solved_velocity_field {
public:
    gridVector2dtype v;
    double t = 0.0;
    void store_velocity_field(const gridVector2dtype& nv, double time) { v = nv; t = time; }
};

// GFM_ITM_interface (external PhD-Interface-tracking-methods)
class GFM_ITM_interface {
public:
    virtual ~GFM_ITM_interface() = default;
    virtual double get_sdf(int i, int j) const = 0;
    virtual void set_sdf(int i, int j, double v) = 0;
    virtual vec2 get_normal(int i, int j) const = 0;
    // at a position, as the reference calls it (GFM_base.cpp:19,
    // GFM_real.cpp:59); every call site passes a cell centre
    virtual vec2 get_normal(const vec2& pos) const = 0;
    virtual bool is_interfacial_cell(int i, int j) const = 0;
    virtual void initialisation() = 0;
    virtual void update_interface_coupled(const solved_velocity_field& vfield, double t, double dt) = 0;
};

// Levelset(params, order, 1) (sim_twofluid.cpp:555-570): HOUC advection of the
// given order with TVD-RK3, reinitialisation every reinit_every steps
class Levelset : public GFM_ITM_interface {
public:
    Levelset(const sim_info& params, int order, int reinit_every = 5, int reinit_iters = 10);
    double get_sdf(int i, int j) const override { return phi[(size_t)i * W + j]; }
    void set_sdf(int i, int j, double v) override { phi[(size_t)i * W + j] = v; }
    vec2 get_normal(int i, int j) const override;
    vec2 get_normal(const vec2& pos) const override;   // the normal of the cell containing pos
    bool is_interfacial_cell(int i, int j) const override;
    void initialisation() override {}
    void update_interface_coupled(const solved_velocity_field& vfield, double t, double dt) override;
    std::vector<double> phi;    // (Ny+2g) x (Nx+2g)
    sim_info params;
    int order, reinit_every, reinit_iters, nsteps = 0;
    int W;
};

// ---------------------------------------------------------------- ghost fluid methods
class GFM_base {
public:
    // GFM_base.hpp:20-21 and the constructor's defaults there; forwarded to
    // the library (gfm_set_gfm_flags) by every set_ghost_states call
    bool use_extension_vfield = true;
    bool use_slip_BCs = false;
    std::shared_ptr<solved_velocity_field> vfield;
    std::shared_ptr<mixed_RS_base> mixed_RS;
    explicit GFM_base(const sim_info&)
        : vfield(std::make_shared<solved_velocity_field>()), mixed_RS(std::make_shared<mixed_RS_exact>()) {}
    virtual ~GFM_base() = default;
    virtual void set_ghost_states(const sim_info& params, const GFM_ITM_interface& ls, const binarySGparams& eosparams,
                                  const double t, grideuler2type& grid1, grideuler2type& grid2, BBrange& realcells1,
                                  BBrange& realcells2) = 0;

protected:
    void run_variant(int variant, const sim_info& params, const GFM_ITM_interface& ls,
                     const binarySGparams& eosparams, double t, grideuler2type& grid1, grideuler2type& grid2,
                     BBrange& realcells1, BBrange& realcells2);
};

#define GFM2D_VARIANT(NAME, ID)                                                                                   \
    class NAME : public GFM_base {                                                                                \
    public:                                                                                                       \
        using GFM_base::GFM_base;                                                                                 \
        void set_ghost_states(const sim_info& params, const GFM_ITM_interface& ls, const binarySGparams& eos,    \
                              const double t, grideuler2type& g1, grideuler2type& g2, BBrange& r1,               \
                              BBrange& r2) override {                                                             \
            run_variant(ID, params, ls, eos, t, g1, g2, r1, r2);                                                  \
        }                                                                                                         \
    };
GFM2D_VARIANT(GFM_original, GFM_ORIGINAL)
GFM2D_VARIANT(GFM_real, GFM_REAL)
GFM2D_VARIANT(GFM_riemann, GFM_RIEMANN)
GFM2D_VARIANT(GFM_modified, GFM_MODIFIED)
// GFM_VOF.hpp / OGFM_VOF.hpp: the VOF-based fills, on the volume fraction, mixed-cell normal and
// interface location of the level set's PLIC reconstruction (the tracker `ls` is a Levelset here)
GFM2D_VARIANT(GFM_VOF, ::GFM_VOFGFM)
GFM2D_VARIANT(OGFM_VOF, ::GFM_OGFMVOF)
#undef GFM2D_VARIANT

// extrapolate_vector.hpp:8-15 (primitive grids)
void extrapolate_vector(const sim_info& params, const GFM_ITM_interface& ls, const int N, grideuler2type& grid1,
                        grideuler2type& grid2);

// BCs of euler_bc.hpp:5-103 on a host grid
void apply_BCs_euler(const sim_info& params, grideuler2type& grid);

// ---------------------------------------------------------------- simulations
class sim_base {
public:
    virtual ~sim_base() = default;
    virtual void run_sim(GFM_settingsfile SF) = 0;
};

class sim_onefluid : public sim_base {
public:
    void run_sim(GFM_settingsfile SF) override;
};

class sim_twofluid : public sim_base {
public:
    void run_sim(GFM_settingsfile SF) override;
};

// parameters of a settings-file test case (sim_twofluid::set_sim_parameters)
void set_twofluid_parameters(const GFM_settingsfile& SF, sim_info& params, binarySGparams& eos);
void set_onefluid_parameters(const GFM_settingsfile& SF, sim_info& params, binarySGparams& eos);

// throws std::runtime_error with the library's message on a non-zero code
void check(int rc, const char* what);

}  // namespace gfm2d
