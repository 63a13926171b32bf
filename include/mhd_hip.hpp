/* C++ front of the C ABI (mhd_hip.h) in the shape of the reference's head/ solver
 * interface, for callers of head/cpu/solver_cpu.hpp:49-72 and
 * head/gpu/solver_gpu.cuh:28-53.  Header-only; link libmhd_hip.so.  A non-OK status
 * throws std::runtime_error with mhd_last_error(), as the reference CPU path throws
 * (solver_cpu.cpp:304-306); the reference GPU path's exit() is not reproduced.
 *
 * One Run is what main_gpu.cu holds per case: the three device grids (Uold, Utmp,
 * Unew), the workspace and the physics constants (gamma, c_h, c_r), which the
 * reference keeps in process globals (head/physics.hpp:26-43) and which are per run
 * here. */
#ifndef MHD_HIP_HPP
#define MHD_HIP_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mhd_hip.h"

namespace mhd_hip {

/* head/riemann.hpp:23 and head/test_cases.hpp:11-23, same numeric orders */
enum class RiemannSolver { HLL = MHD_HLL, HLLC = MHD_HLLC, HLLD = MHD_HLLD, FORCE = MHD_FORCE };
enum class BoundaryType {
    Periodic = MHD_BC_PERIODIC,
    Transmissive = MHD_BC_TRANSMISSIVE,
    Reflecting = MHD_BC_REFLECTING,
    Dirichlet = MHD_BC_DIRICHLET
};
struct BoundaryConfig {
    BoundaryType left = BoundaryType::Transmissive;
    BoundaryType right = BoundaryType::Transmissive;
    BoundaryType bottom = BoundaryType::Transmissive;
    BoundaryType top = BoundaryType::Transmissive;
};
/* The reference limits with minmod only (solver_cpu.cpp:26-43); van Leer is added. */
enum class Limiter { Minmod = MHD_LIMITER_MINMOD, VanLeer = MHD_LIMITER_VANLEER };

inline void check(int rc) {
    if (rc != MHD_OK) throw std::runtime_error(std::string("mhd_hip: ") + mhd_last_error());
}

class Run {
public:
    /* nx, ny interior cells (ng = 2), domain, gamma (init.cpp:20), CFL, solver, BCs. */
    Run(int nx, int ny, double x_min, double x_max, double y_min, double y_max, double gamma, double cfl,
        RiemannSolver solver, const BoundaryConfig& bc, Limiter limiter = Limiter::Minmod, double cr = 0.18,
        void* stream = nullptr)
        : cfl_(cfl) {
        mhd_run_desc d{};
        d.nx = nx;
        d.ny = ny;
        d.ng = 2;
        d.x_min = x_min;
        d.x_max = x_max;
        d.y_min = y_min;
        d.y_max = y_max;
        d.gamma = gamma;
        d.cr = cr;
        d.cfl = cfl;
        d.solver = static_cast<int>(solver);
        d.limiter = static_cast<int>(limiter);
        d.bc[0] = static_cast<int>(bc.left);
        d.bc[1] = static_cast<int>(bc.right);
        d.bc[2] = static_cast<int>(bc.bottom);
        d.bc[3] = static_cast<int>(bc.top);
        d.stream = stream;
        check(mhd_run_create(&d, &run_));
    }
    ~Run() { mhd_run_destroy(run_); }
    Run(const Run&) = delete;
    Run& operator=(const Run&) = delete;
    Run(Run&& o) noexcept : run_(std::exchange(o.run_, nullptr)), cfl_(o.cfl_) {}
    Run& operator=(Run&& o) noexcept {
        std::swap(run_, o.run_);
        std::swap(cfl_, o.cfl_);
        return *this;
    }

    /* Grid2DGPU::upload_from_aos / download_to_aos (grid_gpu.cuh:139-189): reference-order
     * Conserved records, (ny+4) x (nx+4), ghosts included. */
    void upload_from_aos(const double* U) { check(mhd_set_state(run_, U, 0)); }
    void download_to_aos(double* U) { check(mhd_get_state(run_, U, 0)); }
    /* Dense device SoA (9 planes of (nx+4)(ny+4) doubles). */
    void upload_device_soa(const double* U) { check(mhd_set_state(run_, U, 2)); }
    void download_device_soa(double* U) { check(mhd_get_state(run_, U, 2)); }

    /* compute_dt (solver_cpu.hpp:49) / compute_dt_gpu (solver_gpu.cuh:40): ratchets
     * c_h to the maximum signal speed and returns the library's (cfl * min(dx, dy)) / c_h.
     * The run's CFL number is fixed when it is constructed; a different `cfl` throws
     * std::invalid_argument rather than silently computing a second dt. */
    double compute_dt(double cfl) {
        if (!(cfl == cfl_))
            throw std::invalid_argument("mhd_hip: compute_dt(" + std::to_string(cfl) + ") on a run built with CFL " +
                                        std::to_string(cfl_));
        return compute_dt();
    }
    double compute_dt() {
        double dt = 0.0;
        check(mhd_compute_dt(run_, &dt));
        return dt;
    }
    double cfl() const { return cfl_; }
    /* advance_second_order(_gpu) (solver_cpu.hpp:55-63, solver_gpu.cuh:45-53) with the
     * run's solver and BCs, followed by the caller's swap (main_gpu.cu:236). */
    void advance_second_order(double dt) { check(mhd_advance(run_, dt)); }
    /* The grid-view form: caller-owned device grids (dense SoA). */
    void advance_second_order(const double* Uold, double* Utmp, double* Unew, double dt) {
        check(mhd_advance_grids(run_, Uold, Utmp, Unew, dt));
    }
    /* apply_boundary(_gpu) (boundary_cpu.hpp:45, boundary_gpu.cuh:9) on the state. */
    void apply_boundary() { check(mhd_apply_boundary(run_)); }

    /* phys::ch_glm (physics.hpp:41) / set_gpu_physics_ch (solver_gpu.cuh:29) */
    double ch() const {
        double c = 0.0;
        check(mhd_get_ch(run_, &c));
        return c;
    }
    void set_ch(double c) { check(mhd_set_ch(run_, c)); }
    void synchronize() { check(mhd_sync(run_)); }
    /* Queue the run's work on `stream` (a hipStream_t) from now on (mhd_run_set_stream). */
    void set_stream(void* stream) { check(mhd_run_set_stream(run_, stream)); }

    /* initialise_grid + upload_from_aos (init.cpp:34-43) computed on the device, for the
     * cases that have a device state (mhd_run_init_state). */
    void init_state_device(int case_id, const mhd_ic_params* params = nullptr) {
        check(mhd_run_init_state(run_, case_id, params));
    }
    /* write_all_fields (main_cpu.cpp:66-91) of the current state. */
    void write_fields_csv(const std::string& dir, const std::string& prefix) {
        check(mhd_write_fields_csv(run_, dir.c_str(), prefix.c_str()));
    }
    /* Exact dump / restore of the state and c_h. */
    void write_state_bin(const std::string& path, double t, long long step) {
        check(mhd_write_state_bin(run_, path.c_str(), t, step));
    }
    std::pair<double, long long> read_state_bin(const std::string& path) {
        double t = 0.0;
        long long step = 0;
        check(mhd_read_state_bin(run_, path.c_str(), &t, &step));
        return {t, step};
    }
    /* The time loop of main_cpu.cpp:176-276 for case `case_id` (name for the files),
     * with its snapshot schedule; out_dir empty = no files. */
    mhd_drive_result drive(int case_id, const std::string& case_name, double t_end, const std::string& out_dir,
                           int n_scale = 1, double print_interval = 0.1, bool verbose = true) {
        int n = 0;
        check(mhd_case_snapshots(case_id, n_scale, nullptr, nullptr, 0, 0, &n));
        std::vector<double> times(n > 0 ? n : 1);
        std::vector<char> tags(16 * (n > 0 ? n : 1));
        if (n > 0) check(mhd_case_snapshots(case_id, n_scale, times.data(), tags.data(), 16, n, &n));
        std::vector<const char*> tp(n > 0 ? n : 1);
        for (int k = 0; k < n; ++k) tp[k] = tags.data() + 16 * k;
        mhd_drive_opts o{};
        o.case_name = case_name.c_str();
        o.platform = "gpu";
        o.out_dir = out_dir.empty() ? nullptr : out_dir.c_str();
        o.n_scale = n_scale;
        o.t_end = t_end;
        o.n_snapshots = n;
        o.snapshot_times = times.data();
        o.snapshot_tags = tp.data();
        o.print_interval = print_interval;
        o.verbose = verbose ? 1 : 0;
        mhd_drive_result r{};
        check(mhd_drive(run_, &o, &r));
        return r;
    }

    mhd_run* handle() { return run_; }

private:
    mhd_run* run_ = nullptr;
    double cfl_;
};

}  // namespace mhd_hip

#endif /* MHD_HIP_HPP */
