// sfci: command-line driver with the reference's interface (src/main.cpp:58-280).
//
// Reads ./sfci.input (key=value, split at the first '=', no trimming, unknown keys
// ignored; src/main.cpp:128-151), loads the _IN integral file named by `filename`, and
// for task=ground runs the truncated Lanczos solve from init_hf on the GPU; task=excited
// runs sparse_main (src/sparse_solver.cpp:774-809): the CISD state 2 of init_hf, then one
// shift-invert step at target_E with Nl Krylov products per restart.
// Keys: task, Nd, Nl, num_restart, num_orb, num_e, filename, init_hf, coeff_cut,
// target_E, num_thread, bucket_unit (the last two are accepted and unused: the GPU
// path needs neither a thread count nor bucket over-allocation).  Additional keys:
// hpsi_cap (default 60000000), device (default 0), workspace_gb.
// Differences from the reference: coeff_cut is honoured (the reference overwrites it
// with 1e-5, src/main.cpp:196), task=excited runs the call the reference's branch
// leaves commented out (src/main.cpp:235-242), and num_orb up to 128 is accepted for
// task=ground (the reference stops at 64, src/main.cpp:84-85): such a run uses the
// 128-bit-key entry points, with init_hf read as a decimal number of up to 128 bits.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/sfci_gpu.h"

// decimal string -> 128-bit value (lo, hi); false on a non-digit or overflow
static bool parse_u128(const std::string& s, uint64_t& lo, uint64_t& hi) {
    if (s.empty()) return false;
    unsigned __int128 x = 0;
    const unsigned __int128 top = ~(unsigned __int128)0;
    for (char ch : s) {
        if (ch < '0' || ch > '9') return false;
        const unsigned d = (unsigned)(ch - '0');
        if (x > (top - d) / 10) return false;
        x = x * 10 + d;
    }
    lo = (uint64_t)x;
    hi = (uint64_t)(x >> 64);
    return true;
}

// task=ground beyond 64 spin orbitals: the same solve on the 128-bit-key entry points
static int run_wide(const std::string& filename, int num_orb, int num_e, int device, const std::string& init_str,
                    long long Nd, long long Nl, long long num_restart, long long hpsi_cap, double coeff_cut,
                    double workspace_gb) {
    uint64_t init[2] = {0, 0};
    if (!parse_u128(init_str, init[0], init[1])) {
        fprintf(stderr, "bad value for init_hf: %s\n", init_str.c_str());
        return 1;
    }
    sfci_wctx* w = nullptr;
    int rc = sfci_wide_create(filename.c_str(), num_orb, num_e, device, &w);
    if (rc) {
        fprintf(stderr, "failed to load %s (%d)\n", filename.c_str(), rc);
        return 1;
    }
    std::cout << "integrals loaded\n";
    sfci_wide_set_cut(w, coeff_cut);
    if (workspace_gb > 0) sfci_wide_set_workspace(w, (uint64_t)(workspace_gb * 1e9));
    sfci_info_t info;
    sfci_wide_info(w, &info);
    std::cout.precision(12);
    std::cout << "\n*****************parameters summary*****************\n";
    std::cout << "input file:" << filename << "\norbital num:" << num_orb << "\nelectron num:" << num_e
              << "\nNd:" << Nd << "\nstring dim:" << info.string_dim << "\nconnections per det:" << info.n_conn
              << "\nkeys:128-bit\nSpMSpV final result cut:" << coeff_cut << "\n";
    std::cout << "\n****************************************************\n";
    sfci_wvec* hf = nullptr;
    double one = 1.0, ehf = 0.0;
    if ((rc = sfci_wide_vec_create(w, 1, &hf)) || (rc = sfci_wide_vec_upload(hf, init, &one, 1)) ||
        (rc = sfci_wide_energy(w, hf, &ehf))) {
        fprintf(stderr, "init_hf: %s\n", sfci_wide_last_error(w));
        if (hf) sfci_wide_vec_destroy(hf);
        sfci_wide_destroy(w);
        return 1;
    }
    sfci_wide_vec_destroy(hf);
    std::cout << "E_hf:" << ehf << std::endl;
    sfci_lanczos_params_t p{Nd, (int)Nl, (int)num_restart, hpsi_cap, 1e-5, 1, 0.0, nullptr};
    std::vector<double> ritz(num_restart), et(num_restart);
    sfci_lanczos_result_t r;
    rc = sfci_wide_ground_state(w, init, &p, ritz.data(), et.data(), &r, nullptr);
    if (rc) {
        fprintf(stderr, "solve failed (%d): %s\n", rc, sfci_wide_last_error(w));
        sfci_wide_destroy(w);
        return 1;
    }
    std::cout << "SFCI time: " << r.ms_total / 1000.0 << std::endl;
    sfci_wide_destroy(w);
    return 0;
}

int main(int argc, char** argv) {
    std::string input = argc > 1 ? argv[1] : "sfci.input";
    std::string task, filename;
    long long Nd = 200000, Nl = 10, num_restart = 2, hpsi_cap = 60000000;
    int num_orb = 0, num_e = 0, device = 0;
    unsigned long long init_hf = 0;
    std::string init_str = "0";
    double coeff_cut = 1e-5, target_E = 0.0, workspace_gb = 0.0;
    std::ifstream f(input);
    if (!f.is_open()) {
        fprintf(stderr, "cannot open %s\n", input.c_str());
        return 1;
    }
    std::string raw;
    while (std::getline(f, raw)) {
        auto pos = raw.find('=');
        if (pos == std::string::npos) continue;
        std::string key = raw.substr(0, pos), val = raw.substr(pos + 1);
        try {
            if (key == "num_thread" || key == "bucket_unit") {
                std::cout << "using input " << key << "=" << val << " (not used on the GPU)" << std::endl;
                continue;
            } else if (key == "Nd") Nd = std::stoll(val);
            else if (key == "Nl") Nl = std::stoll(val);
            else if (key == "num_restart") num_restart = std::stoll(val);
            else if (key == "num_orb") num_orb = std::stoi(val);
            else if (key == "num_e") num_e = std::stoi(val);
            else if (key == "filename") filename = val;
            else if (key == "coeff_cut") coeff_cut = std::stod(val);
            else if (key == "init_hf") {
                init_str = val;
                uint64_t lo = 0, hi = 0;
                if (!parse_u128(val, lo, hi)) throw std::invalid_argument(val);
                init_hf = lo;
            }
            else if (key == "task") task = val;
            else if (key == "target_E") target_E = std::stod(val);
            else if (key == "hpsi_cap") hpsi_cap = std::stoll(val);
            else if (key == "device") device = std::stoi(val);
            else if (key == "workspace_gb") workspace_gb = std::stod(val);
            else continue;
        } catch (const std::exception&) {
            fprintf(stderr, "bad value for %s: %s\n", key.c_str(), val.c_str());
            return 1;
        }
        std::cout << "using input " << key << "=" << val << std::endl;
    }
    if (num_orb > 128) {
        fprintf(stderr, "orb num larger than 128 is not supported\n");
        return 1;
    }
    {
        uint64_t lo = 0, hi = 0;
        if (num_orb <= 64 && (!parse_u128(init_str, lo, hi) || hi != 0)) {
            fprintf(stderr, "bad value for init_hf: %s\n", init_str.c_str());
            return 1;
        }
    }
    if (num_orb > 64) {
        if (task != "ground") {
            fprintf(stderr, "task=%s is not available beyond 64 spin orbitals (task=ground is)\n", task.c_str());
            return 1;
        }
        return run_wide(filename, num_orb, num_e, device, init_str, Nd, Nl, num_restart, hpsi_cap, coeff_cut,
                        workspace_gb);
    }
    sfci_ctx* ctx = nullptr;
    int rc = sfci_create(filename.c_str(), num_orb, num_e, device, &ctx);
    if (rc) {
        fprintf(stderr, "failed to load %s (%d)\n", filename.c_str(), rc);
        return 1;
    }
    std::cout << "integrals loaded\n";
    sfci_set_cut(ctx, coeff_cut);
    if (workspace_gb > 0) sfci_set_workspace(ctx, (uint64_t)(workspace_gb * 1e9));
    sfci_info_t info;
    sfci_info(ctx, &info);
    std::cout.precision(12);
    std::cout << "\n*****************parameters summary*****************\n";
    std::cout << "input file:" << filename << "\norbital num:" << num_orb << "\nelectron num:" << num_e
              << "\nNd:" << Nd << "\nstring dim:" << info.string_dim << "\nconnections per det:" << info.n_conn
              << "\ntarget energy:" << target_E << "\nSpMSpV final result cut:" << coeff_cut << "\n";
    std::cout << "\n****************************************************\n";
    if (task == "excited") {
        sfci_vec* v = nullptr;
        if ((rc = sfci_vec_create(ctx, Nd, &v))) {
            fprintf(stderr, "vector allocation failed (%d)\n", rc);
            sfci_destroy(ctx);
            return 1;
        }
        sfci_linear_params_t lp{Nd, (int)Nl, (int)num_restart, 1};
        double e = 0.0, e_cisd = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        rc = sfci_excited(ctx, init_hf, target_E, &lp, v, &e, &e_cisd);
        if (rc) {
            fprintf(stderr, "excited-state solve failed (%d): %s\n", rc, sfci_last_error(ctx));
            sfci_vec_destroy(v);
            sfci_destroy(ctx);
            return 1;
        }
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "From CISD: " << e_cisd << "\nenergy after 0th SI:" << e << "\nSFCI time: " << dt << std::endl;
        sfci_vec_destroy(v);
    } else if (task == "ground") {
        double ehf = 0.0;
        if ((rc = sfci_diag_element(ctx, init_hf, &ehf))) {
            fprintf(stderr, "init_hf: %s\n", sfci_last_error(ctx));
            sfci_destroy(ctx);
            return 1;
        }
        std::cout << "E_hf:" << ehf << std::endl;
        sfci_lanczos_params_t p{Nd, (int)Nl, (int)num_restart, hpsi_cap, 1e-5, 1, 0.0};
        std::vector<double> ritz(num_restart), et(num_restart);
        sfci_lanczos_result_t r;
        rc = sfci_ground_state(ctx, init_hf, &p, ritz.data(), et.data(), &r, nullptr);
        if (rc) {
            fprintf(stderr, "solve failed (%d): %s\n", rc, sfci_last_error(ctx));
            sfci_destroy(ctx);
            return 1;
        }
        std::cout << "SFCI time: " << r.ms_total / 1000.0 << std::endl;
    }
    sfci_destroy(ctx);
    return 0;
}
