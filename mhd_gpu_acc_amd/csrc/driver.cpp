// Run driver and output formats of the drop-in interface:
//
//   * the snapshot schedule of the case table (src/test_cases.cpp:206-232, 263-264);
//   * the field CSV writer (scripts/cpu/main_cpu.cpp:36-91): nine files per
//     snapshot, rows = j from the top, "%.8e" values, written to "<path>.tmp" and
//     renamed, so a reader sees a complete file or none;
//   * binary FP64 state dumps (this build's restart / comparison format);
//   * the time loop (main_cpu.cpp:176-276, main_gpu.cu:198-280): dt clamped to the
//     next snapshot or t_end, the non-finite and floor guards, progress lines with
//     max |B|, snapshot writes, the [TIMING] line and the final write.
//
// The loop runs on the device step (mhd_compute_dt + mhd_advance); only the
// snapshot and progress values come back to the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mhd_hip.h"
#include "launch.hpp"
#include "physics.hpp"
#include "run.hpp"

using mhd::fail;

namespace {

constexpr char kMagic[8] = {'M', 'H', 'D', 'S', 'T', 'A', 'T', 'E'};

// Header of a binary state dump: 128 bytes, then 9 planes of (nx+2ng)(ny+2ng)
// little-endian doubles (field order rho, rho u, rho v, rho w, Bx, By, Bz, E, psi;
// each plane row-major with pitch nx+2ng, ghost cells included).  c_h is stored so
// that a run restarted from the dump continues with the same dt sequence.
struct BinHeader {
    char magic[8];
    uint32_t version;       // 1
    uint32_t header_bytes;  // 128
    int32_t nx, ny, ng, nfields;
    double t;
    int64_t step;
    double gamma;
    double ch;
    uint32_t layout;  // 1: SoA planes, ghosts included
    uint32_t reserved[15];
};
static_assert(sizeof(BinHeader) == 128, "dump header is 128 bytes");

// One value as iostream's std::scientific with precision 8 prints it ("%.8e").
inline char* put_value(char* p, char* end, double v) { return std::to_chars(p, end, v, std::chars_format::scientific, 8).ptr; }

// write_field_csv (main_cpu.cpp:36-64) for one field of a SoA grid.
template <class F>
int write_field_csv(const std::string& path, int nx, int ny, F value) {
    const std::string tmp = path + ".tmp";
    std::vector<std::string> rows(ny);
    auto format_rows = [&](int r0, int r1) {
        for (int r = r0; r < r1; ++r) {
            const int j = ny - 1 - r;  // top row first
            std::string& s = rows[r];
            s.resize((size_t)nx * 24 + 1);
            char* p = s.data();
            char* end = p + s.size();
            for (int i = 0; i < nx; ++i) {
                if (i > 0) *p++ = ',';
                p = put_value(p, end, value(i, j));
            }
            *p++ = '\n';
            s.resize(p - s.data());
        }
    };
    const int nt = (long long)nx * ny > 65536 ? std::clamp((int)std::thread::hardware_concurrency(), 1, 16) : 1;
    if (nt == 1) {
        format_rows(0, ny);
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < nt; ++k) pool.emplace_back(format_rows, (int)((long long)ny * k / nt),
                                                       (int)((long long)ny * (k + 1) / nt));
        for (std::thread& th : pool) th.join();
    }
    {
        std::FILE* f = std::fopen(tmp.c_str(), "wb");
        if (!f) return fail(MHD_ERR_IO, "Cannot open output file: " + tmp);
        bool ok = true;
        for (const std::string& s : rows) ok = ok && std::fwrite(s.data(), 1, s.size(), f) == s.size();
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(MHD_ERR_IO, "Write error: " + tmp);
    }
    std::error_code ec;
    std::filesystem::rename(tmp, path, ec);
    if (ec) return fail(MHD_ERR_IO, "rename " + tmp + ": " + ec.message());
    return MHD_OK;
}

// write_all_fields (main_cpu.cpp:66-91) from a SoA grid (9 planes, pitch nx+4).
int write_all_fields(const double* soa, int nx, int ny, double gamma, const std::string& dir,
                     const std::string& prefix) {
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) return fail(MHD_ERR_IO, "create_directories " + dir + ": " + ec.message());
    const int tnx = nx + 4;
    const size_t ps = (size_t)tnx * (ny + 4);
    auto at = [&](int f, int i, int j) { return soa[f * ps + (size_t)(j + 2) * tnx + (i + 2)]; };
    const mhd::Phys ph{gamma, gamma - 1.0, 0.0, 0.0, 0.0, 0.0};
    const auto path = [&](const char* name) { return dir + "/" + prefix + "_" + name + ".csv"; };
    int rc = MHD_OK;
    auto w = [&](const char* name, auto fn) {
        if (rc == MHD_OK) rc = write_field_csv(path(name), nx, ny, fn);
    };
    w("rho", [&](int i, int j) { return at(0, i, j); });
    w("u", [&](int i, int j) { return at(1, i, j) / at(0, i, j); });
    w("v", [&](int i, int j) { return at(2, i, j) / at(0, i, j); });
    w("Bx", [&](int i, int j) { return at(4, i, j); });
    w("By", [&](int i, int j) { return at(5, i, j); });
    w("Bz", [&](int i, int j) { return at(6, i, j); });
    w("E", [&](int i, int j) { return at(7, i, j); });
    w("psi", [&](int i, int j) { return at(8, i, j); });
    w("p", [&](int i, int j) {
        const mhd::Cons u{at(0, i, j), at(1, i, j), at(2, i, j), at(3, i, j), at(4, i, j),
                          at(5, i, j), at(6, i, j), at(7, i, j), at(8, i, j)};
        return mhd::to_prim(u, ph).p;  // phys::cons_to_prim(U).p
    });
    return rc;
}

int download_soa(mhd_run* r, std::vector<double>& soa) {
    soa.resize(9 * r->cells);
    return mhd_get_state(r, soa.data(), 1);
}

// max |B| over the interior of the run's current state (main_cpu.cpp:224-230).
int max_b(mhd_run* r, double* out) {
    MHD_HIP(hipMemsetAsync(r->acc, 0, sizeof(unsigned long long), r->stream()));
    MHD_HIP(mhd::launch_max_b(r->U[0], r->d.nx, r->d.ny, (long long)r->cells, r->acc, r->stream()));
    unsigned long long bits = 0;
    MHD_HIP(hipMemcpyAsync(&bits, r->acc, sizeof bits, hipMemcpyDeviceToHost, r->stream()));
    MHD_HIP(hipStreamSynchronize(r->stream()));
    r->acc_valid = false;  // the accumulator no longer holds the CFL maximum
    std::memcpy(out, &bits, sizeof bits);
    return MHD_OK;
}

}  // namespace

extern "C" {

int mhd_case_snapshots(int case_id, int n_scale, double* times, char* tags, int tag_len, int max_count, int* count) {
    (void)n_scale;  // the schedules do not depend on the grid scale
    if (!count || (max_count > 0 && (!times || !tags || tag_len < 8))) return fail(MHD_ERR_INVALID, "bad arguments");
    std::vector<double> t;
    std::vector<std::string> g;
    if (case_id == MHD_CASE_SHOCK_BUBBLE) {
        // Haas & Sturtevant (1987) Fig. 4 times: T = T0 + t_paper R / vs, with
        // T0 = (xc - R - xs) / vs the shock's arrival at the bubble (test_cases.cpp:206-232)
        constexpr double gam = 1.4, p0 = 1.01325e5, rho = 1.29, ms = 1.22, xc = 0.035, rb = 0.025, xs = 0.005;
        const double vs = ms * std::sqrt(gam * p0 / rho);
        const double t0 = (xc - rb - xs) / vs;
        const double tau = rb / vs;
        for (double tp : {0.6, 1.2, 1.8, 3.0, 4.6, 6.2, 7.8, 12.6, 19.0}) {
            t.push_back(t0 + tp * tau);
            char b[16];
            std::snprintf(b, sizeof b, "t%03d", static_cast<int>(std::round(tp * 10.0)));
            g.emplace_back(b);
        }
    } else if (case_id == MHD_CASE_ORSZAG_TANG) {
        t = {M_PI, 2.0 * M_PI};  // test_cases.cpp:263-264
        g = {"tpi", "t2pi"};
    } else if (case_id < 0 || case_id > MHD_CASE_OT_BUMP) {
        return fail(MHD_ERR_INVALID, "unknown case");
    }
    *count = (int)t.size();
    if (max_count < (int)t.size()) return max_count == 0 ? MHD_OK : fail(MHD_ERR_INVALID, "max_count too small");
    for (size_t k = 0; k < t.size(); ++k) {
        times[k] = t[k];
        std::snprintf(tags + k * tag_len, tag_len, "%s", g[k].c_str());
    }
    return MHD_OK;
}

int mhd_write_fields_csv_host(const double* soa, int nx, int ny, double gamma, const char* dir, const char* prefix) {
    if (!soa || !dir || !prefix || nx < 1 || ny < 1) return fail(MHD_ERR_INVALID, "bad arguments");
    return write_all_fields(soa, nx, ny, gamma, dir, prefix);
}

int mhd_write_fields_csv(mhd_run* r, const char* dir, const char* prefix) {
    if (!r || !dir || !prefix) return fail(MHD_ERR_INVALID, "null argument");
    std::vector<double> soa;
    if (int rc = download_soa(r, soa)) return rc;
    return write_all_fields(soa.data(), r->d.nx, r->d.ny, r->d.gamma, dir, prefix);
}

int mhd_write_state_bin(mhd_run* r, const char* path, double t, long long step) {
    if (!r || !path) return fail(MHD_ERR_INVALID, "null argument");
    std::vector<double> soa;
    if (int rc = download_soa(r, soa)) return rc;
    BinHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.header_bytes = sizeof(BinHeader);
    h.nx = r->d.nx;
    h.ny = r->d.ny;
    h.ng = 2;
    h.nfields = 9;
    h.t = t;
    h.step = step;
    h.gamma = r->d.gamma;
    h.ch = r->ch;
    h.layout = 1;
    const std::string tmp = std::string(path) + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(MHD_ERR_IO, "Cannot open output file: " + tmp);
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(soa.data(), sizeof(double), soa.size(), f) == soa.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(MHD_ERR_IO, "Write error: " + tmp);
    std::error_code ec;
    std::filesystem::rename(tmp, path, ec);
    if (ec) return fail(MHD_ERR_IO, "rename " + tmp + ": " + ec.message());
    return MHD_OK;
}

int mhd_read_state_bin(mhd_run* r, const char* path, double* t, long long* step) {
    if (!r || !path) return fail(MHD_ERR_INVALID, "null argument");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(MHD_ERR_IO, std::string("Cannot open ") + path);
    BinHeader h{};
    std::vector<double> soa(9 * r->cells);
    bool ok = std::fread(&h, sizeof h, 1, f) == 1;
    if (ok && (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1 || h.header_bytes != sizeof h ||
               h.layout != 1 || h.nfields != 9)) {
        std::fclose(f);
        return fail(MHD_ERR_INVALID, std::string(path) + ": not an MHD state dump (version 1)");
    }
    if (ok && (h.nx != r->d.nx || h.ny != r->d.ny || h.ng != 2)) {
        std::fclose(f);
        return fail(MHD_ERR_INVALID, std::string(path) + ": grid " + std::to_string(h.nx) + "x" +
                                         std::to_string(h.ny) + " does not match the run");
    }
    ok = ok && std::fread(soa.data(), sizeof(double), soa.size(), f) == soa.size();
    std::fclose(f);
    if (!ok) return fail(MHD_ERR_IO, std::string(path) + ": short read");
    if (int rc = mhd_set_state(r, soa.data(), 1)) return rc;
    r->ch = h.ch;
    if (t) *t = h.t;
    if (step) *step = h.step;
    return MHD_OK;
}

int mhd_drive(mhd_run* r, const mhd_drive_opts* o, mhd_drive_result* res) {
    if (!r || !o || !res || !o->case_name) return fail(MHD_ERR_INVALID, "null argument");
    if (o->n_snapshots > 0 && (!o->snapshot_times || !o->snapshot_tags))
        return fail(MHD_ERR_INVALID, "snapshot schedule without times or tags");
    std::ofstream logf;
    if (o->log_path) {
        logf.open(o->log_path, std::ios::app);
        if (!logf) return fail(MHD_ERR_IO, std::string("Cannot open log ") + o->log_path);
    }
    std::ostream& out = o->log_path ? static_cast<std::ostream&>(logf) : std::cout;
    const bool write_out = o->out_dir != nullptr;
    const std::string dir = write_out ? o->out_dir : "";
    const std::string base = std::string(o->case_name) + "_" + (o->platform ? o->platform : "gpu") + "_";
    std::vector<double> soa;
    auto write = [&](const std::string& tag) -> int {
        if (int rc = download_soa(r, soa)) return rc;
        if (int rc = write_all_fields(soa.data(), r->d.nx, r->d.ny, r->d.gamma, dir, base + tag)) return rc;
        if (o->verbose) out << "  Wrote fields to " << dir << "/" << base + tag << "_*.csv\n";
        return MHD_OK;
    };
    *res = mhd_drive_result{};
    const bool restart = o->step_start > 0;
    if (write_out && !restart)
        if (int rc = write("t0")) return rc;

    const bool has_snaps = o->n_snapshots > 0;
    int snap = 0;
    double t = restart ? o->t_start : 0.0, t_print_next = t;
    long long step = restart ? o->step_start : 0;
    while (restart && snap < o->n_snapshots && t >= o->snapshot_times[snap] - 1e-12) ++snap;
    int status = MHD_DRIVE_DONE;
    const auto wall0 = std::chrono::steady_clock::now();
    if (o->verbose) {
        out << "\n  step       t         dt        max|B|\n";
        out << "  ----  ----------  ----------  ----------\n";
    }
    int err = MHD_OK;  // an error stops the loop; result still reports the t and step reached
    while (t < o->t_end) {
        double t_next = o->t_end;
        if (has_snaps && snap < o->n_snapshots) t_next = std::min(t_next, o->snapshot_times[snap]);
        double dt_raw = 0.0;
        if ((err = mhd_compute_dt(r, &dt_raw))) break;  // the reference throws (solver_cpu.cpp:304-306)
        const double dt = std::min(dt_raw, t_next - t);
        res->last_dt = dt;
        if (!std::isfinite(dt) || dt <= 0.0) {
            if (o->verbose)
                std::cerr << "[ERROR] dt=" << dt << " at step=" << step << " t=" << t << " (dt_raw=" << dt_raw
                          << "). Aborting.\n";
            status = MHD_DRIVE_BAD_DT;
            break;
        }
        const double dt_floor = 1.0e-8 * o->t_end;
        if (dt < dt_floor) {
            if (o->verbose)
                std::cerr << "[WARN]  dt=" << std::scientific << dt << " (floor=" << dt_floor << ") at step=" << step
                          << " t=" << t << " — numerical blowup, stopping.\n";
            status = MHD_DRIVE_DT_FLOOR;
            break;
        }
        if ((err = mhd_advance(r, dt))) break;
        t += dt;
        step += 1;

        if (o->print_interval > 0.0 && (t >= t_print_next || t >= o->t_end)) {
            double mb = 0.0;
            if ((err = max_b(r, &mb))) break;
            if (o->verbose)
                out << "  " << std::setw(4) << step << "  " << std::setw(10) << std::fixed << std::setprecision(5) << t
                    << "  " << std::setw(10) << std::scientific << std::setprecision(3) << dt << "  "
                    << std::setw(10) << mb << "\n";
            t_print_next = t + o->print_interval;
        }
        // Snapshots reached.  The index advances whether or not files are written,
        // so a run without output goes on to t_end (the reference clamps the next
        // dt to the passed snapshot and stops with dt = 0 there).
        if (has_snaps) {
            while (snap < o->n_snapshots && t >= o->snapshot_times[snap] - 1e-12) {
                if (write_out) {
                    if (o->verbose)
                        out << "  [snap] " << o->snapshot_tags[snap] << "  t_phys=" << std::scientific
                            << std::setprecision(6) << t << " s\n";
                    if ((err = write(o->snapshot_tags[snap]))) break;
                }
                ++snap;
            }
            if (err) break;
        }
        if (o->max_steps > 0 && step >= o->max_steps && t < o->t_end) {
            status = MHD_DRIVE_MAX_STEPS;
            break;
        }
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    res->t = t;
    res->steps = step;
    res->elapsed_s = elapsed;
    res->status = status;
    if (err) return err;
    if (o->verbose) {
        const double sps = static_cast<double>(step) / elapsed;
        const double mcs = static_cast<double>(step) * r->d.nx * r->d.ny / elapsed / 1.0e6;
        out << "\n  Done: " << step << " steps in " << elapsed << " s  (" << sps << " steps/s)\n";
        out << "[TIMING]" << " n=" << o->n_scale << " nx=" << r->d.nx << " ny=" << r->d.ny << " steps=" << step
            << " elapsed_s=" << std::fixed << std::setprecision(4) << elapsed << " steps_per_s=" << std::fixed
            << std::setprecision(2) << sps << " Mcell_updates_s=" << std::fixed << std::setprecision(1) << mcs
            << "\n" << std::flush;
    }
    // final state for cases without a schedule, also after a guard stopped the loop
    // (main_cpu.cpp:272-276); not after a max_steps stop, which the reference lacks
    if (write_out && !has_snaps && status != MHD_DRIVE_MAX_STEPS)
        if (int rc = write("t" + std::to_string(static_cast<int>(std::round(o->t_end * 100))))) return rc;
    return MHD_OK;
}

}  // extern "C"
